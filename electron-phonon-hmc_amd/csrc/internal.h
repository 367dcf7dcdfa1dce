// internal.h -- context and kernel interfaces of libephmc (HIP, gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>
#include <map>
#include <algorithm>
#include "../../include/ephmc.h"

namespace ephmc {

void set_error(const char *fmt, ...);

#define EPHMC_HIP(call)                                                                      \
	do {                                                                                     \
		hipError_t e_ = (call);                                                              \
		if(e_ != hipSuccess){                                                                \
			ephmc::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
			return EPHMC_ERR_HIP;                                                            \
		}                                                                                    \
	} while(0)

#define EPHMC_LAUNCHED()                                                                     \
	do {                                                                                     \
		hipError_t e_ = hipGetLastError();                                                   \
		if(e_ != hipSuccess){                                                                \
			ephmc::set_error("%s:%d launch: %s", __FILE__, __LINE__, hipGetErrorString(e_)); \
			return EPHMC_ERR_HIP;                                                            \
		}                                                                                    \
	} while(0)

#define EPHMC_TRY(expr) do { int r_ = (expr); if(r_ != EPHMC_OK) return r_; } while(0)

// a * b + c rounded once.  The stencil kernels spell every multiply-add with it so that the
// site and strip forms round identically whatever contraction the compiler would choose.
__device__ __forceinline__ double fmac(double a, double b, double c){ return fma(a, b, c); }
__device__ __forceinline__ float fmac(float a, float b, float c){ return fmaf(a, b, c); }

constexpr int MAX_DEG = 8;      // polynomial degree of the slice exponential (2 * pairs)
constexpr int MAX_PAIRS = 4;

// Taylor polynomial of one slice: pair factors in application order and the monomial
// coefficients of their product p(z) = sum_n c[n] z^n.
struct Poly {
	int np;
	double a_lin[MAX_PAIRS], a_quad[MAX_PAIRS];
	int cplx[MAX_PAIRS];
	int deg;
	double c[MAX_DEG + 1];
};

// Kernel-argument bundle of the slice stencil.
template <typename R>
struct Stencil {
	const R *coef;        // [nt][6][ns] bond coefficients t_k - kappa_k x
	const int *nb;        // [6][ns] neighbour sites
	R mu;
	unsigned ns;
};

// Kernel-group timing: event pairs recorded on the context stream, read back on demand.
struct TimerGroup {
	std::vector<hipEvent_t> ev;   // begin/end pairs
	size_t used;                  // events recorded since the last harvest
	bool open;
	double total_ms;
	unsigned long calls;
};

// Transient-localisation workspace for a batch of cap_b disorder samples (allocated on first use).
struct TlWork {
	unsigned cap_b = 0, cap_nbins = 0;
	size_t rot_cap = 0;               // rotations recorded per sample
	unsigned sw_cap = 0;              // sweeps recorded per sample
	double *A = nullptr, *Z = nullptr;   // [B][N][N]; the sorted eigenvectors V reuse A
	double2 *rot = nullptr;           // [B][rot_cap] (c, s); W_0, W_1 ([B][N][N] each) reuse it
	double *V = nullptr, *W0 = nullptr, *W1 = nullptr;
	int4 *sw = nullptr;               // [B][sw_cap] (first rotation, lowest, highest position)
	unsigned *nsw = nullptr;
	int *status = nullptr, *perm = nullptr;
	double *x = nullptr, *d = nullptr, *e = nullptr, *tau = nullptr, *ev = nullptr, *evs = nullptr, *ems = nullptr;
	double *res = nullptr, *part = nullptr, *acc = nullptr, *red = nullptr;
};

// Buffers of the Schur-form inverse (SCHUR_FORM_INV; allocated on first use): the quasi-triangular
// factor, the Schur vectors, a product, reflector factors, eigenvalues, the threshold record and
// the iteration status.
struct SchurWork {
	double *H = nullptr, *Z = nullptr, *P = nullptr, *tau = nullptr, *wr = nullptr, *wi = nullptr, *info = nullptr;
	int *status = nullptr;
};

// Buffers of the stabilised slice product (allocated on first use): U, V, the QR work matrix and
// X (N x N each), the column scales d, reflector factors tau, the column order, the reflector count.
struct StabWork {
	double *U = nullptr, *V = nullptr, *ZR = nullptr, *X = nullptr, *d = nullptr, *tau = nullptr;
	int *perm = nullptr, *nref = nullptr;
};

} // namespace ephmc

struct ephmc_ctx {
	ephmc_config cfg;
	unsigned l1, l2, ns, nt, loc;     // loc = 3 ns links per slice
	size_t dim, mat;                  // dim = nt*loc, mat = ns*ns
	double traj_length;
	ephmc::Poly poly;
	hipStream_t stream;
	bool own_stream;
	int device;

	std::vector<unsigned> nnt;        // host copy of the lattice table (12 ns)
	int *d_nb;                        // [6][ns]
	unsigned *d_link;                 // [6][ns]

	// slice coefficients
	double *d_coef64;                 // [nt][6][ns]
	float *d_coef32;

	// chain / force workspace
	unsigned ckpt;                    // k: T stored every k slices (k = 1: every slice, read back directly)
	unsigned nseg;                    // ceil(nt / k) stored matrices
	unsigned span;                    // slices per launch of the force sweep (k, or 32 when k = 1)
	unsigned cols;                    // columns per workgroup of the force sweep
	unsigned nwg;                     // ceil(ns / cols)
	// forward-chain launch shape chosen by ephmc_autotune, per arithmetic ([0] fp64, [1] fp32):
	// columns per workgroup (0: rule) and sites per thread of a lattice row (-1: rule, 0: one)
	unsigned chain_cols_set[2];
	int chain_strip_set[2];
	int adj_w0_early;                 // force sweep: W_0 of a slice written ahead (one barrier fewer per slice)
	size_t partial_cap;               // doubles in each d_partial buffer
	void *d_ck;                       // checkpoints T_{s k}: nseg matrices (R = fp64 or fp32)
	void *d_seg;                      // per-workgroup segment scratch of the recompute: (ns + 3) * ns * k values
	double *d_partial[2];             // [span][nwg][3 ns] force partial sums, double-buffered
	double *d_Tfin;                   // T_nt (fp64) or scratch for conversions
	float *d_Tfin32;

	// Gauss-Jordan
	double *d_gj[2];
	int *d_piv, *d_src;
 	double *d_diag;
	double *d_dlu;                    // D^{-1}, L of the current / next diagonal block: 4 x (32 x 32)
	double *d_scal;                   // device scalars (see SCAL_* below)

	// bosons
	double *d_twiddle;                // [nt/2] complex e^{-2 pi i k/nt}
	double *d_f2;                     // [nt] freq2(n) with omega
	double *d_f2mom;                  // [nt] kinetic frequencies: d_f2, or with sqrt(cfa_mass^2 + omega^2) (CFA)
	double *d_red;                    // reduction partials
	unsigned red_blocks;

	// correlator (allocated on first use): R_t0, R_0t for every slice, g_ab(tau)
	double *d_Rt0, *d_R0t, *d_ccg;

	// transient localisation
	ephmc::TlWork tl;

	// stabilised slice product (ephmc_config.stabilize, ephmc_stable_logdet_inv)
	ephmc::StabWork stab;

	// Schur-form inverse (ephmc_config.schur_form_inv, ephmc_schur)
	ephmc::SchurWork schur;

	// chain state
	double *d_x, *d_p, *d_pdot, *d_xnew, *d_minv, *d_mhat, *d_normals;
	double *h_stage[2];               // pinned staging for normals
	hipEvent_t stage_ev[2];
	int stage_next;
	bool chain_ready;
	unsigned batch_refs;              // batches holding this context's buffers (ephmc_batch_create)

	// captured trajectory body (ephmc_config.graph)
	hipGraph_t graph;
	hipGraphExec_t graph_exec;
	bool graph_failed;

	// timing
	bool timing;
	std::map<std::string, ephmc::TimerGroup> timers;
};

namespace ephmc {

enum Scalar {            // layout of d_scal
	SCAL_LOGDET = 0,      // log|det| of the last factorisation
	SCAL_SIGN = 1,
	SCAL_EP = 2,          // kinetic energy
	SCAL_H_OLD = 3,
	SCAL_H_NEW = 4,
	SCAL_LOGDET_CHAIN = 5,// log det of the accepted state
	SCAL_ACC = 6,
	SCAL_BOLTZ = 7,
	SCAL_ENERGY = 8,
	SCAL_DH = 9,
	SCAL_LOGDET_NEW = 10,
	SCAL_BOSON = 11,      // S_B(x)
	SCAL_GJ_FLAG = 12,    // a multiplier above 1 was seen by the Gauss-Jordan fast path
	SCAL_GJ_FALLBACKS = 13, // number of pivoted fallback inversions run
	SCAL_FSCALE = 14,     // force factor exp(-logdet) of SINGLE_PARTICLE
	SCAL_SINGULAR = 15,   // 1 + step of a zero pivot column in the last inversion, or 0
	SCAL_SCHUR_FAIL = 16, // 1 + index where the Schur iteration of the last inversion stopped, or 0
	SCAL_COUNT = 20
};

// One launch over many chains (batch.hip): workgroup (x, y, z) is block (x, y) of chain z, whose
// argument record sits at index z of a device table; a chain's blocks past its own grid return.
// The per-chain records are collected here on the host and uploaded once.
struct BatchOp {
	const void *fn = nullptr;
	unsigned block = 0, gx = 0, gy = 1;
	size_t lds = 0, rec = 0, off = 0;     // dynamic LDS (largest chain), record bytes, table offset
	unsigned n = 0;                       // chains in this launch
	std::vector<unsigned char> tab;
	template <typename A>
	int add(const void *f, unsigned b, size_t l, unsigned x, unsigned y, const A &a)
	{
		if(fn && (fn != f || block != b || rec != sizeof(A))){
			set_error("batch: one launch would mix two kernels");
			return EPHMC_ERR_ARG;
		}
		fn = f;
		block = b;
		rec = sizeof(A);
		lds = std::max(lds, l);
		gx = std::max(gx, x);
		gy = std::max(gy, y);
		const unsigned char *p = reinterpret_cast<const unsigned char *>(&a);
		tab.insert(tab.end(), p, p + sizeof(A));
		n++;
		return EPHMC_OK;
	}
};

// timing helpers
int timer_begin(ephmc_ctx *c, const char *group);
int timer_end(ephmc_ctx *c, const char *group);

// chain.hip
unsigned sites_per_thread(const char *var);
unsigned strip_width(const ephmc_ctx *c, int prec);
int coefficients(ephmc_ctx *c, const double *x, bool want32);
int chain_forward(ephmc_ctx *c, int prec, const void *src, void *dst, void *ck, unsigned ck_every,
		int descending, bool from_identity);
int chain_range(ephmc_ctx *c, int prec, const double *src, void *dst, int tau0, int dir, int ntau);
int apply_exp_partial(ephmc_ctx *c, const double *x_tau, double *M, int revert, int stop);
int slice_coefficients(ephmc_ctx *c, const double *x_tau);
int transpose_plus_identity(ephmc_ctx *c, int prec, const void *T, double *M, int add_identity);
int convert_f64_to_f32(ephmc_ctx *c, const double *a, float *b, size_t n);
int add_identity(ephmc_ctx *c, double *M);
int log_trace(ephmc_ctx *c, int prec, const void *M, int ld_slot, int fs_slot);
int set_identity(ephmc_ctx *c, int prec, void *M);
int set_scalar(ephmc_ctx *c, int slot, double v);
// batch records (threads: workgroup size of the site-form kernels' LDS layout, 256 / 512 / 1024;
// block: threads launched, a multiple of 64 with ns <= block <= threads)
int batch_coef(BatchOp &op, ephmc_ctx *c, const double *x, bool want32);
int batch_chain(BatchOp &op, ephmc_ctx *c, int prec, unsigned threads, unsigned block, void *dst, void *ck,
		unsigned ck_every, int descending);
int batch_transpose(BatchOp &op, ephmc_ctx *c, int prec, const void *T, double *M);
int batch_add_identity(BatchOp &op, ephmc_ctx *c, double *M);
// dst <- src (n doubles); with gate, only when gate[SCAL_ACC] != 0
int batch_copy(BatchOp &op, const double *src, double *dst, size_t n, const double *gate);

// gj.hip
int gj_logdet_inverse(ephmc_ctx *c, double *M);
// the same inversion as batch records: first block, steps[i] = i-th block step, pivoted fallback
int batch_gj(BatchOp &first, std::vector<BatchOp> &steps, BatchOp &fallback, ephmc_ctx *c, double *M);

// adjoint.hip
int force_adjoint(ephmc_ctx *c, int prec, const double *minv, double *pdot, double kick, double *p, const double *fscale);
int force_constant(ephmc_ctx *c, int prec, const double *minv, double *pdot, const double *fscale);
int broadcast_slice0(ephmc_ctx *c, double *pdot);
// LDS bytes of one force-sweep workgroup at C columns (the launch shape used for ns sites)
size_t adjoint_lds(unsigned ns, int prec, unsigned C);
// columns per workgroup of the force sweep (the context's choice, else 2 in fp64 and 4 in fp32)
unsigned adjoint_cols(const ephmc_ctx *c, int prec);
// ctx.hip: sets the context's force-sweep columns per workgroup (reserving its partial sums)
int set_force_cols(ephmc_ctx *c, unsigned C);
// force_adjoint(c, prec, minv, pdot, 0, null, null) as batch records: segs[i] = i-th sweep launch,
// then the final reduction
int batch_adjoint(std::vector<BatchOp> &segs, BatchOp &reduce, ephmc_ctx *c, int prec, unsigned threads,
		unsigned block, const double *minv, double *pdot);

// bosons.hip
int fa_evolve(ephmc_ctx *c, double *x, double *p, double h, const double *pdot, double kick);
int fa_filter(ephmc_ctx *c, const double *in, double *out, int inverse_sqrt);
int fa_energy(ephmc_ctx *c, const double *p, double *scal_out);
int fa_drift(ephmc_ctx *c, const double *p, double *x, double h);
int boson_force(ephmc_ctx *c, const double *x, double *pdot, int overwrite);
int axpy(ephmc_ctx *c, double *y, double a, const double *x, size_t n);
int plain_energy(ephmc_ctx *c, const double *p, double *scal_out);
int boson_action(ephmc_ctx *c, const double *x, double *scal_out);
int hamilton_dev(ephmc_ctx *c, const double *x, const double *p, int logdet_slot, int out_slot);
int kinetic_energy(ephmc_ctx *c, const double *p, double *scal_out);
int metropolis(ephmc_ctx *c, double uniform);
// batch records of fa_evolve, fa_filter (momentum refresh), hamilton_dev (five launches, ops[0..4]) and of
// the Metropolis decision with the uniform read from device memory
int batch_fa_evolve(BatchOp &op, ephmc_ctx *c, double *x, double *p, double h, const double *pdot, double kick);
int batch_fa_filter(BatchOp &op, ephmc_ctx *c, const double *in, double *out);
int batch_hamilton(BatchOp *ops, ephmc_ctx *c, const double *x, const double *p, int logdet_slot, int out_slot);
int batch_metropolis(BatchOp &op, ephmc_ctx *c, const double *uniform);

// tl.hip
void tl_release(ephmc_ctx *c);

// stab.hip
int stab_reserve(ephmc_ctx *c);
void stab_release(ephmc_ctx *c);
int stable_logdet_inverse(ephmc_ctx *c, int prec, unsigned k, double *minv);

// schur.hip
int schur_reserve(ephmc_ctx *c);
void schur_release(ephmc_ctx *c);
int schur_logdet_inverse(ephmc_ctx *c, double *M);
int schur_constant_force(ephmc_ctx *c, const double *minv, double *pdot);

} // namespace ephmc
