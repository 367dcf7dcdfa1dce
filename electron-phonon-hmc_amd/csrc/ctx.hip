// ctx.hip -- context, memory, C-ABI entry points and the device-resident HMC driver.
#include "internal.h"
#include <functional>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <cstdlib>
#include <vector>

namespace ephmc {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_err, sizeof(g_err), fmt, ap);
	va_end(ap);
}

unsigned chain_cols(const ephmc_ctx *c, int prec);

// Taylor pair coefficients of the truncated exponential series (the tables of
// taylor-coeff_pairs.h for the even orders; their pair products equal sum_{k<=order} z^k/k!)
// and the application order of set_taylor_pair_order (organic_dense_matrix.c:133-159).
static int make_poly(int order, Poly &P)
{
	static const double t2[][2] = {{1, 1}};
	static const double t4[][2] = {{1.8294933739891903, 0.9404019519941707}, {0.1705066260108097, 1.5785318126884535}};
	static const double t6[][2] = {{2.2561549994606459, 0.8008468192975424}, {1.0805822170351833, 1.8245938269821296},
		{-0.3367372164958292, 1.5494473211106643}};
	static const double t8[][2] = {{2.5115385528762511, 0.6852625793459637}, {1.6810015288012661, 1.7481130845728929},
		{0.4249750401763088, 2.0321227226190966}, {-0.6175151218538260, 1.4284655522676216}};
	const double (*tab)[2];
	int np;
	switch(order){
		case 2: tab = t2; np = 1; break;
		case 4: tab = t4; np = 2; break;
		case 6: tab = t6; np = 3; break;
		case 8: tab = t8; np = 4; break;
		default: return EPHMC_ERR_ARG;
	}
	int ord[MAX_PAIRS];
	const int tuples = np / 2;
	int lo = 0, hi = tuples - 1;
	for(int q = 0; q < tuples; q++){
		const int tup = (q % 3 == 1) ? lo++ : hi--;
		ord[2 * q] = tup;
		ord[2 * q + 1] = np - 1 - tup;
	}
	if(np % 2) ord[2 * tuples] = tuples;
	P.np = np;
	double poly[MAX_DEG + 1] = {1};
	int deg = 0;
	for(int i = 0; i < np; i++){
		const double re = .5 * tab[ord[i]][0] / np, im = .5 * tab[ord[i]][1] / np;
		P.a_lin[i] = 2 * re;
		P.a_quad[i] = std::pow(std::hypot(re, im), 2);   // |a|^2 as pow(cabs(a), 2), organic_dense_matrix.c:169
		P.cplx[i] = im != 0;
		const double f[3] = {1, P.a_lin[i], P.cplx[i] ? P.a_quad[i] : 0};
		const int fd = P.cplx[i] ? 2 : 1;
		double nxt[MAX_DEG + 1] = {0};
		for(int a = 0; a <= deg; a++) for(int b = 0; b <= fd; b++) nxt[a + b] += poly[a] * f[b];
		deg += fd;
		std::copy(nxt, nxt + MAX_DEG + 1, poly);
	}
	P.deg = deg;
	std::copy(poly, poly + MAX_DEG + 1, P.c);
	return EPHMC_OK;
}

// Periodic triangular parallelogram, organic_lattice.c:10-46 (same table layout).
static void make_lattice(unsigned l1, unsigned l2, std::vector<unsigned> &nnt)
{
	const unsigned ns = l1 * l2;
	nnt.assign(12 * (size_t)ns, 0);
	for(unsigned i = 0; i < ns; i++){
		const unsigned x1 = i % l1, x2 = i / l1;
		const unsigned up = (x2 + 1) % l2, dn = (x2 + l2 - 1) % l2;
		unsigned *n = &nnt[6 * i], *k = &nnt[6 * ((size_t)ns + i)];
		n[0] = x2 * l1 + (x1 + 1) % l1;
		n[1] = x2 * l1 + (x1 + l1 - 1) % l1;
		n[2] = up * l1 + x1;
		n[3] = dn * l1 + x1;
		n[4] = up * l1 + (x1 + l1 - 1) % l1;
		n[5] = dn * l1 + (x1 + 1) % l1;
		k[0] = 3 * i;        k[1] = 3 * n[1];
		k[2] = 3 * i + 1;    k[3] = 3 * n[3] + 1;
		k[4] = 3 * i + 2;    k[5] = 3 * n[5] + 2;
	}
}

template <typename T>
static int dalloc(T **p, size_t n)
{
	*p = nullptr;
	if(n == 0) return EPHMC_OK;
	if(hipMalloc((void **)p, n * sizeof(T)) != hipSuccess){
		(void)hipGetLastError();
		set_error("hipMalloc of %zu bytes failed", n * sizeof(T));
		return EPHMC_ERR_ALLOC;
	}
	return EPHMC_OK;
}

static int timer_harvest(TimerGroup &g)
{
	for(size_t i = 0; i + 1 < g.used; i += 2){
		float ms = 0;
		EPHMC_HIP(hipEventSynchronize(g.ev[i + 1]));
		EPHMC_HIP(hipEventElapsedTime(&ms, g.ev[i], g.ev[i + 1]));
		g.total_ms += ms;
		g.calls++;
	}
	g.used = 0;
	return EPHMC_OK;
}

int timer_begin(ephmc_ctx *c, const char *group)
{
	if(!c->timing) return EPHMC_OK;
	TimerGroup &g = c->timers[group];
	if(g.used + 2 > 4096) EPHMC_TRY(timer_harvest(g));
	while(g.ev.size() < g.used + 2){
		hipEvent_t e;
		EPHMC_HIP(hipEventCreate(&e));
		g.ev.push_back(e);
	}
	EPHMC_HIP(hipEventRecord(g.ev[g.used], c->stream));
	g.open = true;
	return EPHMC_OK;
}

int timer_end(ephmc_ctx *c, const char *group)
{
	if(!c->timing) return EPHMC_OK;
	TimerGroup &g = c->timers[group];
	if(!g.open) return EPHMC_OK;
	EPHMC_HIP(hipEventRecord(g.ev[g.used + 1], c->stream));
	g.used += 2;
	g.open = false;
	return EPHMC_OK;
}

// Columns per workgroup of the adjoint sweep: the widest that fits LDS and keeps at least one
// workgroup per CU (at 32x32: 4 columns, 0.13 ms per 32-slice launch in fp32 and 0.32 ms in
// fp64).  EPHMC_ADJOINT_COLS overrides it for tuning sweeps.
static unsigned adjoint_cols_auto(unsigned ns, int prec)
{
	if(const char *e = getenv("EPHMC_ADJOINT_COLS")){
		const unsigned v = (unsigned)atoi(e);
		if(v == 1 || v == 2 || v == 4) return v;
	}
	for(unsigned C = 4; C > 1; C >>= 1)
		if(adjoint_lds(ns, prec, C) <= 160 * 1024 && (ns + C - 1) / C >= 256) return C;
	return 1;
}

// log|det M| and M <- M^{-1}: the Gauss-Jordan kernels, or through the real Schur form of M - 1
// under SCHUR_FORM_INV (organic_dense_matrix.c:556-603)
static int logdet_inverse(ephmc_ctx *c, double *M)
{
	return c->cfg.schur_form_inv ? schur_logdet_inverse(c, M) : gj_logdet_inverse(c, M);
}

// Adjoint force for the current coefficients and checkpoints.  minv: M_hat^{-1} (fp64), or
// null for SINGLE_PARTICLE (the sweep starts from the identity).
static int force_dispatch(ephmc_ctx *c, int prec, const double *minv, double *pdot, const double *fscale)
{
	// CONSTANT_FORCE under SCHUR_FORM_INV: QM1P = 1 - M_hat^{-1}, no chain rule (organic_dense_matrix.c:270-278)
	if(c->cfg.constant_force && c->cfg.schur_form_inv) return schur_constant_force(c, minv, pdot);
	if(c->cfg.constant_force) return force_constant(c, prec, minv, pdot, fscale);
	if(!minv) EPHMC_TRY(set_identity(c, prec, prec == EPHMC_FP64 ? (void *)c->d_Tfin : (void *)c->d_Tfin32));
	return force_adjoint(c, prec, minv, pdot, 0, nullptr, fscale);
}

// ---- one MD force evaluation at field x: coefficients, forward chain with checkpoints,
// M_hat, Gauss-Jordan inverse, adjoint sweep (organic_hmc.c:178-183 + force :86-134).
// SINGLE_PARTICLE: no +1 and no inverse; the sweep starts from the identity and the force is
// scaled by exp(-log tr M_hat).  CONSTANT_FORCE: the slice-0 force on every slice.
// NO_FOURIER_ACC: the boson force is added (organic_hmc.c:131).
static int md_force(ephmc_ctx *c, const double *x, double *pdot)
{
	const int prec = c->cfg.precision;
	const bool sp = c->cfg.single_particle != 0;
	EPHMC_TRY(timer_begin(c, "chain"));
	EPHMC_TRY(coefficients(c, x, prec != EPHMC_FP64));
	void *tfin = prec == EPHMC_FP64 ? (void *)c->d_Tfin : (void *)c->d_Tfin32;
	EPHMC_TRY(chain_forward(c, prec, nullptr, tfin, c->d_ck, c->ckpt, 0, true));
	if(!sp && !c->cfg.stabilize) EPHMC_TRY(transpose_plus_identity(c, prec, tfin, c->d_mhat, 1));
	EPHMC_TRY(timer_end(c, "chain"));
	const double *minv = c->d_mhat, *fscale = nullptr;
	if(c->cfg.stabilize){
		// M_hat^{-1} from the product stabilised every k slices (the slice precision of the molecular
		// dynamics); the sweep below reads the checkpoints of the chain above
		EPHMC_TRY(timer_begin(c, "gj"));
		EPHMC_TRY(stable_logdet_inverse(c, prec, c->cfg.stabilize, c->d_mhat));
		EPHMC_TRY(timer_end(c, "gj"));
	}else if(sp){
		EPHMC_TRY(log_trace(c, prec, tfin, SCAL_LOGDET, SCAL_FSCALE));
		fscale = c->d_scal + SCAL_FSCALE;
		minv = nullptr;
	}else{
		EPHMC_TRY(timer_begin(c, "gj"));
		EPHMC_TRY(logdet_inverse(c, c->d_mhat));
		EPHMC_TRY(timer_end(c, "gj"));
	}
	EPHMC_TRY(timer_begin(c, "adjoint"));
	EPHMC_TRY(force_dispatch(c, prec, minv, pdot, fscale));
	EPHMC_TRY(timer_end(c, "adjoint"));
	if(c->cfg.no_fourier_acc) EPHMC_TRY(boson_force(c, x, pdot, 0));
	return EPHMC_OK;
}

// FP64 M_hat and log det of field x (Metropolis quality), slices applied from nt-1 down to 0 as
// in construct_M_hat (organic_dense_matrix.c:256-258): d_mhat = M_hat^{-1} and the log det in
// scal[SCAL_LOGDET]; SINGLE_PARTICLE: d_mhat = B..B and log tr; ZERO_FILLING: d_mhat = B..B.
static int exact_logdet(ephmc_ctx *c, const double *x)
{
	if(c->cfg.stabilize){
		EPHMC_TRY(coefficients(c, x, false));
		return stable_logdet_inverse(c, EPHMC_FP64, c->cfg.stabilize, c->d_mhat);
	}
	EPHMC_TRY(coefficients(c, x, false));
	EPHMC_TRY(chain_forward(c, EPHMC_FP64, nullptr, c->d_mhat, nullptr, 0, 1, true));
	if(c->cfg.zero_filling) return set_scalar(c, SCAL_LOGDET, 0.0);
	if(c->cfg.single_particle) return log_trace(c, EPHMC_FP64, c->d_mhat, SCAL_LOGDET, -1);
	EPHMC_TRY(add_identity(c, c->d_mhat));
	EPHMC_TRY(logdet_inverse(c, c->d_mhat));
	return EPHMC_OK;
}

// update_x (organic_hmc.c:136-165) for the modes without the fused Fourier-accelerated drift
static int update_x_dev(ephmc_ctx *c, double *x, double *p, double h, double half)
{
	if(c->cfg.no_fourier_acc) return axpy(c, x, h * fabs(half), p, c->dim);
	// CLASSICAL_FOURIER_ACC: cfa_steps approximate sub-steps with the boson force
	const unsigned cs = c->cfg.cfa_steps;
	const double hh = h / cs;
	if(half != .5){
		EPHMC_TRY(boson_force(c, x, c->d_pdot, 1));
		EPHMC_TRY(axpy(c, p, .5 * hh, c->d_pdot, c->dim));
		for(unsigned k = 0; k < cs / 2; k++){
			EPHMC_TRY(fa_drift(c, p, x, hh));
			EPHMC_TRY(boson_force(c, x, c->d_pdot, 1));
			EPHMC_TRY(axpy(c, p, hh, c->d_pdot, c->dim));
		}
		EPHMC_TRY(fa_drift(c, p, x, .5 * hh));
	}
	if(half != -.5){
		EPHMC_TRY(fa_drift(c, p, x, .5 * hh));
		for(unsigned k = 0; k < cs / 2; k++){
			EPHMC_TRY(boson_force(c, x, c->d_pdot, 1));
			EPHMC_TRY(axpy(c, p, hh, c->d_pdot, c->dim));
			EPHMC_TRY(fa_drift(c, p, x, hh));
		}
		EPHMC_TRY(boson_force(c, x, c->d_pdot, 1));
		EPHMC_TRY(axpy(c, p, .5 * hh, c->d_pdot, c->dim));
	}
	return EPHMC_OK;
}

// organic_hmc.c:167-190 on (x, p): x(h/2), then [force, p += h F, x(h)] with the final drift h/2.
static int leap_frog_dev(ephmc_ctx *c, double *x, double *p)
{
	const unsigned steps = c->cfg.steps;
	const double h = steps ? c->traj_length / steps : 2 * c->traj_length;
	if(c->cfg.no_fourier_acc || c->cfg.classical_fourier_acc){
		EPHMC_TRY(update_x_dev(c, x, p, h, .5));
		for(unsigned s = 1; s <= steps; s++){
			EPHMC_TRY(md_force(c, x, c->d_pdot));
			EPHMC_TRY(axpy(c, p, h, c->d_pdot, c->dim));
			EPHMC_TRY(update_x_dev(c, x, p, h, s == steps ? -.5 : 1.0));
		}
		return EPHMC_OK;
	}
	EPHMC_TRY(timer_begin(c, "fft"));
	EPHMC_TRY(fa_evolve(c, x, p, 0.5 * h, nullptr, 0));
	EPHMC_TRY(timer_end(c, "fft"));
	for(unsigned s = 1; s <= steps; s++){
		EPHMC_TRY(md_force(c, x, c->d_pdot));
		const double drift = (s == steps) ? 0.5 * h : h;
		// kick p += h F fused into the drift's load
		EPHMC_TRY(timer_begin(c, "fft"));
		EPHMC_TRY(fa_evolve(c, x, p, drift, c->d_pdot, h));
		EPHMC_TRY(timer_end(c, "fft"));
	}
	return EPHMC_OK;
}

// momenta from dim normals (sample_fourier_momenta, organic_bosons.c:76-100)
static int momenta(ephmc_ctx *c, const double *normals, double *p)
{
	if(c->cfg.no_fourier_acc){
		EPHMC_HIP(hipMemcpyAsync(p, normals, c->dim * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
		return EPHMC_OK;
	}
	return fa_filter(c, normals, p, 0);
}

static int upload_normals(ephmc_ctx *c, const double *normals)
{
	// double-buffered pinned staging so the caller may reuse its buffer on return
	const int s = c->stage_next;
	c->stage_next ^= 1;
	EPHMC_HIP(hipEventSynchronize(c->stage_ev[s]));
	std::memcpy(c->h_stage[s], normals, c->dim * sizeof(double));
	EPHMC_HIP(hipMemcpyAsync(c->d_normals, c->h_stage[s], c->dim * sizeof(double), hipMemcpyHostToDevice, c->stream));
	EPHMC_HIP(hipEventRecord(c->stage_ev[s], c->stream));
	return EPHMC_OK;
}

static int read_scal(ephmc_ctx *c, int slot, double *out)
{
	EPHMC_HIP(hipMemcpyAsync(out, c->d_scal + slot, sizeof(double), hipMemcpyDeviceToHost, c->stream));
	EPHMC_HIP(hipStreamSynchronize(c->stream));
	return EPHMC_OK;
}

} // namespace ephmc

using namespace ephmc;

// =============================================================================================
extern "C" {

const char *ephmc_last_error(void){ return g_err; }

int ephmc_create(const ephmc_config *cfg, ephmc_ctx **out)
{
	if(!cfg || !out){ set_error("null argument"); return EPHMC_ERR_ARG; }
	*out = nullptr;
	if(cfg->l1 == 0 || cfg->l2 == 0 || cfg->nt == 0){ set_error("empty lattice"); return EPHMC_ERR_ARG; }
	// f_n = omega + (2 sin(pi n / nt))^2 / omega (organic_bosons.c:137) divides by omega (negative
	// under static disorder, where the units scale by -beta: organic_main.c:50-53)
	if(!(fabs(cfg->omega) > 0)){ set_error("omega must be nonzero, got %g", cfg->omega); return EPHMC_ERR_ARG; }
	if(cfg->precision != EPHMC_FP64 && cfg->precision != EPHMC_MIXED){ set_error("precision"); return EPHMC_ERR_ARG; }
	if(cfg->classical_fourier_acc && cfg->cfa_steps == 0){ set_error("CLASSICAL_FOURIER_ACC needs cfa_steps >= 1"); return EPHMC_ERR_ARG; }
	if(cfg->no_fourier_acc && cfg->classical_fourier_acc){ set_error("NO_FOURIER_ACC and CLASSICAL_FOURIER_ACC exclude each other"); return EPHMC_ERR_ARG; }
	if(cfg->schur_form_inv && cfg->stabilize){
		set_error("SCHUR_FORM_INV and stabilize exclude each other");
		return EPHMC_ERR_UNSUPPORTED;
	}
	if(cfg->schur_form_inv && cfg->single_particle && cfg->constant_force){
		set_error("SCHUR_FORM_INV with SINGLE_PARTICLE and CONSTANT_FORCE is not provided");
		return EPHMC_ERR_UNSUPPORTED;
	}
	if(cfg->stabilize && (cfg->zero_filling || cfg->single_particle)){
		set_error("stabilize: the stabilised product forms 1 + B..B, not provided for ZERO_FILLING / SINGLE_PARTICLE");
		return EPHMC_ERR_UNSUPPORTED;
	}
	if((unsigned long long)cfg->l1 * cfg->l2 > 2048){
		set_error("lattice of %llu sites exceeds 2048", (unsigned long long)cfg->l1 * cfg->l2);
		return EPHMC_ERR_UNSUPPORTED;
	}
	const unsigned ns = cfg->l1 * cfg->l2;
	ephmc_ctx *c = new ephmc_ctx();
	c->cfg = *cfg;
	c->chain_strip_set[0] = c->chain_strip_set[1] = -1;
	c->l1 = cfg->l1; c->l2 = cfg->l2; c->ns = ns; c->nt = cfg->nt; c->loc = 3 * ns;
	c->dim = (size_t)c->nt * c->loc;
	c->mat = (size_t)ns * ns;
	c->traj_length = cfg->traj_length == 0 ? 1.570796326794897 : cfg->traj_length;
	if(c->cfg.taylor_order == 0) c->cfg.taylor_order = 4;
	if(c->cfg.flavors == 0) c->cfg.flavors = 1;
	int rc = make_poly(c->cfg.taylor_order, c->poly);
	if(rc != EPHMC_OK){ set_error("taylor order %d not provided (2, 4, 6, 8)", cfg->taylor_order); delete c; return rc; }
	c->device = cfg->device;
	if(hipSetDevice(c->device) != hipSuccess){ set_error("hipSetDevice(%d) failed", c->device); delete c; return EPHMC_ERR_HIP; }
	if(cfg->stream){ c->stream = (hipStream_t)cfg->stream; c->own_stream = false; }
	else{
		if(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess){ set_error("stream"); delete c; return EPHMC_ERR_HIP; }
		c->own_stream = true;
	}
	make_lattice(c->l1, c->l2, c->nnt);
	const size_t rsz = cfg->precision == EPHMC_FP64 ? 8 : 4;
	// checkpoint interval: every slice while the Nt partial products fit in 4 GiB (the force
	// sweep then reads T_tau back instead of recomputing it), else 16
	if(cfg->checkpoint) c->ckpt = std::min(cfg->checkpoint, c->nt);
	else c->ckpt = (size_t)c->nt * c->mat * rsz <= ((size_t)4 << 30) ? 1u : std::min(16u, c->nt);
	c->nseg = (c->nt + c->ckpt - 1) / c->ckpt;
	c->span = c->ckpt == 1 ? std::min(32u, c->nt) : c->ckpt;
	c->cols = adjoint_cols_auto(ns, cfg->precision);
	c->adj_w0_early = 1;
	if(const char *e = getenv("EPHMC_ADJOINT_W0EARLY")) c->adj_w0_early = atoi(e) != 0;
	c->nwg = (ns + c->cols - 1) / c->cols;
	c->red_blocks = (unsigned)std::min<size_t>((c->dim + 255) / 256, 1024);
	const size_t fft_blocks = (c->loc + 31) / 32 + 1;

	auto fail = [&](int r){ ephmc_destroy(c); return r; };
	std::vector<int> nb(6 * (size_t)ns);
	std::vector<unsigned> link(6 * (size_t)ns);
	for(unsigned i = 0; i < ns; i++)
		for(int d = 0; d < 6; d++){
			nb[(size_t)d * ns + i] = (int)c->nnt[6 * i + d];
			link[(size_t)d * ns + i] = c->nnt[6 * ((size_t)ns + i) + d];
		}
	const bool mixed = cfg->precision == EPHMC_MIXED;
	if((rc = dalloc(&c->d_nb, nb.size())) || (rc = dalloc(&c->d_link, link.size()))
		|| (rc = dalloc(&c->d_coef64, (size_t)c->nt * 6 * ns))
		|| (rc = dalloc(&c->d_coef32, mixed ? (size_t)c->nt * 6 * ns : 0))
		|| (rc = dalloc((unsigned char **)&c->d_ck, (size_t)c->nseg * c->mat * rsz))
		// recompute scratch of every width 1, 2, 4: ceil(ns / C) C <= ns + 3 columns
		|| (rc = dalloc((unsigned char **)&c->d_seg, ((size_t)ns + 3) * ns * c->ckpt * rsz))
		|| (rc = dalloc(&c->d_partial[0], c->partial_cap = (size_t)c->span * c->nwg * c->loc))
		|| (rc = dalloc(&c->d_partial[1], (size_t)c->span * c->nwg * c->loc))
		|| (rc = dalloc(&c->d_Tfin, c->mat))
		|| (rc = dalloc(&c->d_Tfin32, mixed ? c->mat : 0))
		|| (rc = dalloc(&c->d_gj[0], c->mat)) || (rc = dalloc(&c->d_gj[1], c->mat))
		|| (rc = dalloc(&c->d_piv, ns)) || (rc = dalloc(&c->d_src, ns)) || (rc = dalloc(&c->d_diag, ns))
		|| (rc = dalloc(&c->d_dlu, 4 * 32 * 32))
		|| (rc = dalloc(&c->d_scal, SCAL_COUNT))
		|| (rc = dalloc(&c->d_twiddle, 2 * (size_t)c->nt)) || (rc = dalloc(&c->d_f2, c->nt)) || (rc = dalloc(&c->d_f2mom, c->nt))
		|| (rc = dalloc(&c->d_red, std::max<size_t>(2 * c->red_blocks, fft_blocks)))
		|| (rc = dalloc(&c->d_x, c->dim)) || (rc = dalloc(&c->d_p, c->dim)) || (rc = dalloc(&c->d_pdot, c->dim))
		|| (rc = dalloc(&c->d_xnew, c->dim)) || (rc = dalloc(&c->d_normals, c->dim))
		|| (rc = dalloc(&c->d_minv, c->mat)) || (rc = dalloc(&c->d_mhat, c->mat)))
		return fail(rc);
	for(int s = 0; s < 2; s++){
		if(hipHostMalloc((void **)&c->h_stage[s], c->dim * sizeof(double), hipHostMallocDefault) != hipSuccess){
			set_error("pinned staging allocation failed");
			return fail(EPHMC_ERR_ALLOC);
		}
		if(hipEventCreateWithFlags(&c->stage_ev[s], hipEventDisableTiming) != hipSuccess){ set_error("event"); return fail(EPHMC_ERR_HIP); }
		(void)hipEventRecord(c->stage_ev[s], c->stream);
	}
	std::vector<double> tw(2 * (size_t)c->nt), f2(c->nt), f2m(c->nt);
	const double pi = 3.141592653589793;
	// kinetic frequencies: omega -> sqrt(cfa_mass^2 + omega^2) under CLASSICAL_FOURIER_ACC
	// (organic_hmc.c:61-62, organic_bosons.c:83-84,162-163)
	const double om_k = cfg->classical_fourier_acc ? sqrt(pow(cfg->cfa_mass, 2) + pow(cfg->omega, 2)) : cfg->omega;
	for(unsigned j = 0; j < c->nt; j++){
		tw[2 * j] = cos(2 * pi * j / c->nt);
		tw[2 * j + 1] = -sin(2 * pi * j / c->nt);
		const double s = 2 * sin(pi / c->nt * j);
		f2[j] = cfg->omega + s * s / cfg->omega;   // organic_bosons.c:137
		f2m[j] = om_k + s * s / om_k;
	}
	if(hipMemcpy(c->d_nb, nb.data(), nb.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess
		|| hipMemcpy(c->d_link, link.data(), link.size() * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess
		|| hipMemcpy(c->d_twiddle, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess
		|| hipMemcpy(c->d_f2, f2.data(), f2.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess
		|| hipMemcpy(c->d_f2mom, f2m.data(), f2m.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess
		|| hipMemset(c->d_scal, 0, SCAL_COUNT * sizeof(double)) != hipSuccess){
		set_error("initial uploads failed");
		return fail(EPHMC_ERR_HIP);
	}
	// the stabilised product's buffers now, so that a captured trajectory allocates nothing
	if(cfg->stabilize && (rc = stab_reserve(c)) != EPHMC_OK) return fail(rc);
	if(cfg->schur_form_inv && (rc = schur_reserve(c)) != EPHMC_OK) return fail(rc);
	*out = c;
	return EPHMC_OK;
}

int ephmc_destroy(ephmc_ctx *c)
{
	if(!c) return EPHMC_OK;
	if(c->batch_refs){ set_error("the context belongs to a batch: destroy the batch first"); return EPHMC_ERR_ARG; }
	if(c->stream) (void)hipStreamSynchronize(c->stream);
	void *bufs[] = {c->d_nb, c->d_link, c->d_coef64, c->d_coef32, c->d_ck, c->d_seg, c->d_partial[0], c->d_partial[1], c->d_Tfin,
		c->d_Tfin32, c->d_gj[0], c->d_gj[1], c->d_piv, c->d_src, c->d_diag, c->d_dlu, c->d_scal, c->d_twiddle, c->d_f2, c->d_f2mom, c->d_red,
		c->d_x, c->d_p, c->d_pdot, c->d_xnew, c->d_normals, c->d_minv, c->d_mhat, c->d_Rt0, c->d_R0t, c->d_ccg};
	for(void *b : bufs) if(b) (void)hipFree(b);
	tl_release(c);
	stab_release(c);
	schur_release(c);
	if(c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
	if(c->graph) (void)hipGraphDestroy(c->graph);
	for(int s = 0; s < 2; s++){
		if(c->h_stage[s]) (void)hipHostFree(c->h_stage[s]);
		if(c->stage_ev[s]) (void)hipEventDestroy(c->stage_ev[s]);
	}
	for(auto &kv : c->timers) for(hipEvent_t e : kv.second.ev) (void)hipEventDestroy(e);
	if(c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
	delete c;
	return EPHMC_OK;
}

int ephmc_synchronize(ephmc_ctx *c)
{
	if(!c){ set_error("null context"); return EPHMC_ERR_ARG; }
	EPHMC_HIP(hipStreamSynchronize(c->stream));
	return EPHMC_OK;
}

int ephmc_dims(const ephmc_ctx *c, unsigned *ns, unsigned *nt, size_t *dim)
{
	if(!c){ set_error("null context"); return EPHMC_ERR_ARG; }
	if(ns) *ns = c->ns;
	if(nt) *nt = c->nt;
	if(dim) *dim = c->dim;
	return EPHMC_OK;
}

int ephmc_lattice(const ephmc_ctx *c, unsigned *nnt)
{
	if(!c || !nnt){ set_error("null argument"); return EPHMC_ERR_ARG; }
	std::memcpy(nnt, c->nnt.data(), c->nnt.size() * sizeof(unsigned));
	return EPHMC_OK;
}

int ephmc_malloc(size_t bytes, void **dev)
{
	if(!dev){ set_error("null argument"); return EPHMC_ERR_ARG; }
	if(hipMalloc(dev, bytes) != hipSuccess){ (void)hipGetLastError(); set_error("hipMalloc(%zu) failed", bytes); return EPHMC_ERR_ALLOC; }
	return EPHMC_OK;
}

int ephmc_free(void *dev)
{
	if(dev) EPHMC_HIP(hipFree(dev));
	return EPHMC_OK;
}

int ephmc_memcpy_h2d(ephmc_ctx *c, void *dev, const void *host, size_t bytes)
{
	if(!c){ set_error("null context"); return EPHMC_ERR_ARG; }
	EPHMC_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, c->stream));
	EPHMC_HIP(hipStreamSynchronize(c->stream));
	return EPHMC_OK;
}

int ephmc_memcpy_d2h(ephmc_ctx *c, void *host, const void *dev, size_t bytes)
{
	if(!c){ set_error("null context"); return EPHMC_ERR_ARG; }
	EPHMC_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->stream));
	EPHMC_HIP(hipStreamSynchronize(c->stream));
	return EPHMC_OK;
}

int ephmc_apply_exp_M(ephmc_ctx *c, const double *x_tau, double *M, int revert, int stop)
{
	if(!c || !x_tau || !M){ set_error("null argument"); return EPHMC_ERR_ARG; }
	return apply_exp_partial(c, x_tau, M, revert, stop);
}

int ephmc_construct_M_hat(ephmc_ctx *c, const double *x, double *M_hat)
{
	if(!c || !x || !M_hat){ set_error("null argument"); return EPHMC_ERR_ARG; }
	EPHMC_TRY(coefficients(c, x, false));
	// B_0 (B_1 ( ... B_{nt-1} I)) in the reference's order (slices applied from nt-1 down to 0);
	// no identity for ZERO_FILLING / SINGLE_PARTICLE (organic_dense_matrix.c:260-264)
	EPHMC_TRY(chain_forward(c, EPHMC_FP64, nullptr, M_hat, nullptr, 0, 1, true));
	if(c->cfg.zero_filling || c->cfg.single_particle) return EPHMC_OK;
	return add_identity(c, M_hat);
}

int ephmc_logdet_inv(ephmc_ctx *c, double *M_hat, double *logdet, int *sign)
{
	if(!c || !M_hat){ set_error("null argument"); return EPHMC_ERR_ARG; }
	if(c->cfg.single_particle){
		// log tr M_hat, M_hat unchanged (organic_dense_matrix.c:550-551)
		EPHMC_TRY(log_trace(c, EPHMC_FP64, M_hat, SCAL_LOGDET, -1));
		if(logdet) EPHMC_TRY(read_scal(c, SCAL_LOGDET, logdet));
		if(sign) *sign = 1;
		return EPHMC_OK;
	}
	EPHMC_TRY(timer_begin(c, "gj"));
	EPHMC_TRY(logdet_inverse(c, M_hat));
	EPHMC_TRY(timer_end(c, "gj"));
	if(!logdet && !sign) return EPHMC_OK;
	double ld, sg, sing, sfail;
	EPHMC_TRY(read_scal(c, SCAL_LOGDET, &ld));
	EPHMC_TRY(read_scal(c, SCAL_SIGN, &sg));
	EPHMC_TRY(read_scal(c, SCAL_SINGULAR, &sing));
	EPHMC_TRY(read_scal(c, SCAL_SCHUR_FAIL, &sfail));
	if(logdet) *logdet = ld;
	if(sfail != 0){
		set_error("Schur form: the QR iteration did not converge at index %d", (int)sfail - 1);
		return EPHMC_ERR_NUMERIC;
	}
	if(sign) *sign = sg < 0 ? -1 : 1;
	// a zero pivot: log|det| = -inf as in the reference (organic_dense_matrix.c:582-584)
	if(sing != 0 || ld == -HUGE_VAL){
		if(sing != 0) set_error("singular matrix: column %d has no nonzero pivot; log|det| = -inf", (int)sing);
		else set_error("singular matrix: zero pivot; log|det| = -inf");
		return EPHMC_ERR_NUMERIC;
	}
	return EPHMC_OK;
}

int ephmc_last_pivots(ephmc_ctx *c, int *ipiv)
{
	if(!c || !ipiv){ set_error("null argument"); return EPHMC_ERR_ARG; }
	EPHMC_HIP(hipMemcpyAsync(ipiv, c->d_piv, c->ns * sizeof(int), hipMemcpyDeviceToHost, c->stream));
	EPHMC_HIP(hipStreamSynchronize(c->stream));
	for(unsigned i = 0; i < c->ns; i++) ipiv[i] += 1;        // LAPACK numbering
	return EPHMC_OK;
}

int ephmc_force(ephmc_ctx *c, const double *x, const double *M_inv, double logdet, double *p_dot)
{
	if(!c || !x || !p_dot || (!M_inv && !c->cfg.single_particle)){ set_error("null argument"); return EPHMC_ERR_ARG; }
	const int prec = c->cfg.precision;
	EPHMC_TRY(coefficients(c, x, prec != EPHMC_FP64));
	void *tfin = prec == EPHMC_FP64 ? (void *)c->d_Tfin : (void *)c->d_Tfin32;
	EPHMC_TRY(timer_begin(c, "chain"));
	EPHMC_TRY(chain_forward(c, prec, nullptr, tfin, c->d_ck, c->ckpt, 0, true));
	EPHMC_TRY(timer_end(c, "chain"));
	const double *fscale = nullptr;
	if(c->cfg.single_particle){
		EPHMC_TRY(set_scalar(c, SCAL_FSCALE, exp(-logdet)));
		fscale = c->d_scal + SCAL_FSCALE;
		M_inv = nullptr;
	}
	EPHMC_TRY(timer_begin(c, "adjoint"));
	EPHMC_TRY(force_dispatch(c, prec, M_inv, p_dot, fscale));
	EPHMC_TRY(timer_end(c, "adjoint"));
	if(c->cfg.no_fourier_acc) EPHMC_TRY(boson_force(c, x, p_dot, 0));
	return EPHMC_OK;
}

int ephmc_md_force(ephmc_ctx *c, const double *x, double *p_dot)
{
	if(!c || !x || !p_dot){ set_error("null argument"); return EPHMC_ERR_ARG; }
	return md_force(c, x, p_dot);
}

int ephmc_sample_free_bosons(ephmc_ctx *c, const double *normals, double *x)
{
	if(!c || !normals || !x){ set_error("null argument"); return EPHMC_ERR_ARG; }
	return fa_filter(c, normals, x, 1);
}

int ephmc_sample_fourier_momenta(ephmc_ctx *c, const double *normals, double *p)
{
	if(!c || !normals || !p){ set_error("null argument"); return EPHMC_ERR_ARG; }
	return momenta(c, normals, p);
}

int ephmc_energy_momenta(ephmc_ctx *c, const double *p, double *energy)
{
	if(!c || !p || !energy){ set_error("null argument"); return EPHMC_ERR_ARG; }
	EPHMC_TRY(kinetic_energy(c, p, c->d_scal + SCAL_EP));
	return read_scal(c, SCAL_EP, energy);
}

int ephmc_evolve_bosons(ephmc_ctx *c, double *x, double *p, double h)
{
	if(!c || !x || !p){ set_error("null argument"); return EPHMC_ERR_ARG; }
	return fa_evolve(c, x, p, h, nullptr, 0);
}

int ephmc_update_x(ephmc_ctx *c, double *x, double *p, double h, double half)
{
	if(!c || !x || !p){ set_error("null argument"); return EPHMC_ERR_ARG; }
	if(c->cfg.no_fourier_acc || c->cfg.classical_fourier_acc) return update_x_dev(c, x, p, h, half);
	return fa_evolve(c, x, p, fabs(half) * h, nullptr, 0);
}

int ephmc_hamilton(ephmc_ctx *c, const double *x, double logdet, double energyP, double *H)
{
	if(!c || !x || !H){ set_error("null argument"); return EPHMC_ERR_ARG; }
	EPHMC_TRY(boson_action(c, x, c->d_scal + SCAL_BOSON));
	double sb;
	EPHMC_TRY(read_scal(c, SCAL_BOSON, &sb));
	*H = -logdet * c->cfg.flavors + energyP + sb;
	return EPHMC_OK;
}

int ephmc_leap_frog(ephmc_ctx *c, double *x, double *p)
{
	if(!c || !x || !p){ set_error("null argument"); return EPHMC_ERR_ARG; }
	return leap_frog_dev(c, x, p);
}

int ephmc_chain_init(ephmc_ctx *c, const double *normals)
{
	if(!c || !normals){ set_error("null argument"); return EPHMC_ERR_ARG; }
	EPHMC_TRY(upload_normals(c, normals));
	EPHMC_TRY(fa_filter(c, c->d_normals, c->d_x, 1));
	EPHMC_TRY(exact_logdet(c, c->d_x));
	EPHMC_HIP(hipMemcpyAsync(c->d_minv, c->d_mhat, c->mat * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
	EPHMC_HIP(hipMemcpyAsync(c->d_scal + SCAL_LOGDET_CHAIN, c->d_scal + SCAL_LOGDET, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
	c->chain_ready = true;
	return EPHMC_OK;
}

int ephmc_chain_set_x(ephmc_ctx *c, const double *x_host)
{
	if(!c || !x_host){ set_error("null argument"); return EPHMC_ERR_ARG; }
	EPHMC_HIP(hipMemcpyAsync(c->d_x, x_host, c->dim * sizeof(double), hipMemcpyHostToDevice, c->stream));
	EPHMC_TRY(exact_logdet(c, c->d_x));
	EPHMC_HIP(hipMemcpyAsync(c->d_minv, c->d_mhat, c->mat * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
	EPHMC_HIP(hipMemcpyAsync(c->d_scal + SCAL_LOGDET_CHAIN, c->d_scal + SCAL_LOGDET, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
	EPHMC_HIP(hipStreamSynchronize(c->stream));
	c->chain_ready = true;
	return EPHMC_OK;
}

// Everything of a trajectory between the momentum draw and the Metropolis decision (fixed
// device pointers, no host values), so that it can be captured once and replayed.
static int trajectory_body(ephmc_ctx *c)
{
	EPHMC_HIP(hipMemcpyAsync(c->d_xnew, c->d_x, c->dim * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
	EPHMC_TRY(momenta(c, c->d_normals, c->d_p));
	EPHMC_TRY(hamilton_dev(c, c->d_xnew, c->d_p, SCAL_LOGDET_CHAIN, SCAL_H_OLD));
	EPHMC_TRY(leap_frog_dev(c, c->d_xnew, c->d_p));
	EPHMC_TRY(timer_begin(c, "metropolis"));
	EPHMC_TRY(exact_logdet(c, c->d_xnew));
	EPHMC_TRY(hamilton_dev(c, c->d_xnew, c->d_p, SCAL_LOGDET, SCAL_H_NEW));
	return EPHMC_OK;
}

// Releases a captured trajectory graph (it is captured again on the next trajectory).
static void drop_graph(ephmc_ctx *c)
{
	if(c->graph_exec){
		(void)hipGraphExecDestroy(c->graph_exec);
		c->graph_exec = nullptr;
	}
	if(c->graph){
		(void)hipGraphDestroy(c->graph);
		c->graph = nullptr;
	}
}

// Captures trajectory_body into a hipGraph.  A capture that fails leaves the context on the
// eager path (graph_failed), with the reason in ephmc_last_error.
static int capture_body(ephmc_ctx *c)
{
	if(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess){
		(void)hipGetLastError();
		c->graph_failed = true;
		return EPHMC_OK;
	}
	const int rc = trajectory_body(c);
	hipGraph_t g = nullptr;
	const hipError_t e = hipStreamEndCapture(c->stream, &g);
	if(rc != EPHMC_OK || e != hipSuccess || !g){
		(void)hipGetLastError();
		if(g) (void)hipGraphDestroy(g);
		c->graph_failed = true;
		set_error("graph capture of the trajectory failed; running eagerly");
		return EPHMC_OK;
	}
	if(hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0) != hipSuccess){
		(void)hipGetLastError();
		(void)hipGraphDestroy(g);
		c->graph_exec = nullptr;
		c->graph_failed = true;
		set_error("graph instantiation failed; running eagerly");
		return EPHMC_OK;
	}
	c->graph = g;
	return EPHMC_OK;
}

int ephmc_trajectory(ephmc_ctx *c, const double *normals, double uniform, ephmc_traj_record *rec)
{
	if(!c || !normals){ set_error("null argument"); return EPHMC_ERR_ARG; }
	if(!c->chain_ready){ set_error("ephmc_chain_init has not run"); return EPHMC_ERR_ARG; }
	EPHMC_TRY(upload_normals(c, normals));
	if(c->cfg.zero_filling){
		// organic_hmc.c:246-254: a fresh free-boson field every trajectory, M_hat = B..B, no
		// Metropolis step (the normals are sample_free_bosons' draw; no uniform is used)
		EPHMC_TRY(fa_filter(c, c->d_normals, c->d_x, 1));
		EPHMC_TRY(exact_logdet(c, c->d_x));
		EPHMC_HIP(hipMemcpyAsync(c->d_minv, c->d_mhat, c->mat * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
		EPHMC_TRY(set_scalar(c, SCAL_LOGDET_CHAIN, 0.0));
		EPHMC_TRY(hamilton_dev(c, c->d_x, nullptr, SCAL_LOGDET_CHAIN, SCAL_ENERGY));
		EPHMC_TRY(set_scalar(c, SCAL_ACC, 1.0));
		EPHMC_TRY(set_scalar(c, SCAL_BOLTZ, 1.0));
		EPHMC_TRY(set_scalar(c, SCAL_DH, 0.0));
		if(rec) return ephmc_last_record(c, rec);
		return EPHMC_OK;
	}
	// organic_hmc.c:213-243
	if(c->cfg.graph && !c->timing && !c->graph_failed){
		if(!c->graph_exec) EPHMC_TRY(capture_body(c));
		if(c->graph_exec) EPHMC_HIP(hipGraphLaunch(c->graph_exec, c->stream));
		else EPHMC_TRY(trajectory_body(c));
	}else{
		EPHMC_TRY(trajectory_body(c));
	}
	EPHMC_TRY(metropolis(c, uniform));
	EPHMC_TRY(timer_end(c, "metropolis"));
	if(rec) return ephmc_last_record(c, rec);
	return EPHMC_OK;
}

int ephmc_last_record(ephmc_ctx *c, ephmc_traj_record *rec)
{
	if(!c || !rec){ set_error("null argument"); return EPHMC_ERR_ARG; }
	double s[SCAL_COUNT];
	EPHMC_HIP(hipMemcpyAsync(s, c->d_scal, sizeof(s), hipMemcpyDeviceToHost, c->stream));
	EPHMC_HIP(hipStreamSynchronize(c->stream));
	rec->accepted = s[SCAL_ACC] != 0;
	rec->boltzmann = s[SCAL_BOLTZ];
	rec->energy = s[SCAL_ENERGY];
	rec->logdet = s[SCAL_LOGDET_CHAIN];
	rec->dH = s[SCAL_DH];
	return EPHMC_OK;
}

int ephmc_chain_get_x(ephmc_ctx *c, double *x)
{
	if(!c || !x){ set_error("null argument"); return EPHMC_ERR_ARG; }
	return ephmc_memcpy_d2h(c, x, c->d_x, c->dim * sizeof(double));
}

int ephmc_chain_get_minv(ephmc_ctx *c, double *m)
{
	if(!c || !m){ set_error("null argument"); return EPHMC_ERR_ARG; }
	return ephmc_memcpy_d2h(c, m, c->d_minv, c->mat * sizeof(double));
}

int ephmc_chain_get_logdet(ephmc_ctx *c, double *ld)
{
	if(!c || !ld){ set_error("null argument"); return EPHMC_ERR_ARG; }
	return read_scal(c, SCAL_LOGDET_CHAIN, ld);
}

int ephmc_chain_buffers(ephmc_ctx *c, double **x, double **p, double **minv)
{
	if(!c){ set_error("null context"); return EPHMC_ERR_ARG; }
	if(x) *x = c->d_x;
	if(p) *p = c->d_p;
	if(minv) *minv = c->d_minv;
	return EPHMC_OK;
}

int ephmc_gj_fallbacks(ephmc_ctx *c, unsigned long *count)
{
	if(!c || !count){ set_error("null argument"); return EPHMC_ERR_ARG; }
	double v;
	EPHMC_TRY(read_scal(c, SCAL_GJ_FALLBACKS, &v));
	*count = (unsigned long)v;
	return EPHMC_OK;
}

} // extern "C"

// ---- autotuner ----------------------------------------------------------------------------
// Times the forward chain (coefficients + chain with the context's stored products) for every
// columns-per-workgroup width that fits LDS, then the force sweep for every width, on a zero
// field with M_hat^{-1} = 1 (the kernels do the same work for any values), and keeps the
// fastest of each.  The sweep's partial buffers grow when a narrower width needs more
// workgroups.  A captured trajectory graph is dropped, to be captured again with the new
// shapes.  Uses the MD scratch buffers only; call it between trajectories.
__global__ void zero_kernel(double *a, size_t n)
{
	for(size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = 0;
}

static int set_zero(ephmc_ctx *c, double *a, size_t n)
{
	zero_kernel<<<(unsigned)std::min<size_t>((n + 255) / 256, 1024), 256, 0, c->stream>>>(a, n);
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

// fastest of reps single runs after one warm-up run (code load): a transient slowdown, such as
// the clock still rising, costs one sample rather than the average
static int time_ms(ephmc_ctx *c, hipEvent_t e0, hipEvent_t e1, int reps, double *ms, const std::function<int()> &body)
{
	EPHMC_TRY(body());
	double best = 1e30;
	for(int r = 0; r < reps; r++){
		EPHMC_HIP(hipEventRecord(e0, c->stream));
		EPHMC_TRY(body());
		EPHMC_HIP(hipEventRecord(e1, c->stream));
		EPHMC_HIP(hipEventSynchronize(e1));
		float f = 0;
		EPHMC_HIP(hipEventElapsedTime(&f, e0, e1));
		if(f < best) best = f;
	}
	*ms = best;
	return EPHMC_OK;
}

static int reserve_partials(ephmc_ctx *c, unsigned nwg)
{
	const size_t need = (size_t)c->span * nwg * c->loc;
	if(need <= c->partial_cap) return EPHMC_OK;
	// the new pair first: a failed allocation leaves the context on its old buffers
	double *nb[2] = {nullptr, nullptr};
	if(dalloc(&nb[0], need) != EPHMC_OK || dalloc(&nb[1], need) != EPHMC_OK){
		if(nb[0]) (void)hipFree(nb[0]);
		return EPHMC_ERR_ALLOC;
	}
	EPHMC_HIP(hipStreamSynchronize(c->stream));
	for(int b = 0; b < 2; b++){
		(void)hipFree(c->d_partial[b]);
		c->d_partial[b] = nb[b];
	}
	c->partial_cap = need;
	return EPHMC_OK;
}

// Columns per workgroup of the force sweep set to C (partial sums reserved for its workgroup count;
// a captured trajectory graph, which holds the old shape, is dropped).
int ephmc::set_force_cols(ephmc_ctx *c, unsigned C)
{
	if(c->cols == C) return EPHMC_OK;
	const unsigned nwg = (c->ns + C - 1) / C;
	EPHMC_TRY(reserve_partials(c, nwg));
	drop_graph(c);
	c->cols = C;
	c->nwg = nwg;
	return EPHMC_OK;
}

extern "C" {

int ephmc_autotune(ephmc_ctx *c, ephmc_tune_result *res)
{
	if(!c){ set_error("null context"); return EPHMC_ERR_ARG; }
	// a batch holds the context's buffers and launch shapes, which tuning may replace
	if(c->batch_refs){ set_error("the context belongs to a batch: tune it before ephmc_batch_create"); return EPHMC_ERR_ARG; }
	const int prec = c->cfg.precision;
	const unsigned ns = c->ns;
	EPHMC_HIP(hipStreamSynchronize(c->stream));
	// a captured trajectory graph holds the current buffers and launch shapes: drop it before
	// either can change (it is captured again on the next trajectory)
	drop_graph(c);
	// the shapes in force on entry, restored if tuning fails part way
	const unsigned cols0 = c->cols, nwg0 = c->nwg;
	const int early0 = c->adj_w0_early;
	const unsigned ccols0[2] = {c->chain_cols_set[0], c->chain_cols_set[1]};
	const int cstrip0[2] = {c->chain_strip_set[0], c->chain_strip_set[1]};
	const bool timing = c->timing;
	c->timing = false;
	hipEvent_t e0, e1;
	EPHMC_HIP(hipEventCreate(&e0));
	EPHMC_HIP(hipEventCreate(&e1));
	int rc = EPHMC_OK;
	ephmc_tune_result r{};
	void *tfin = prec == EPHMC_FP64 ? (void *)c->d_Tfin : (void *)c->d_Tfin32;
	// the molecular-dynamics chain (context precision, stores its checkpoints) and, in mixed
	// precision, the fp64 chain of the Metropolis log det (no checkpoints)
	auto md_chain = [&]() -> int {
		EPHMC_TRY(coefficients(c, c->d_xnew, prec != EPHMC_FP64));
		return chain_forward(c, prec, nullptr, tfin, c->d_ck, c->ckpt, 0, true);
	};
	auto ld_chain = [&]() -> int {
		EPHMC_TRY(coefficients(c, c->d_xnew, false));
		return chain_forward(c, EPHMC_FP64, nullptr, c->d_Tfin, nullptr, 0, 1, true);
	};
	// widths: 2 C N values of LDS; strips of M sites need l1 % M == 0 and N / M <= 1024 threads
	auto tune_chain = [&](int cprec, auto &&run, unsigned &cols_out, unsigned &strip_out, double &ms_out) -> int {
		const int slot = cprec == EPHMC_FP64 ? 0 : 1;
		const size_t csz = cprec == EPHMC_FP64 ? 8 : 4;
		struct Shape { unsigned C, M; double ms; };
		std::vector<Shape> shapes;
		for(unsigned M : {0u, 2u, 4u}){
			if(M && (c->l1 % M != 0 || ns / M > 1024)) continue;
			for(unsigned C : {1u, 2u, 4u, 8u}){
				if(2 * (size_t)C * ns * csz > 160 * 1024 || C > ns) continue;
				shapes.push_back({C, M, 1e30});
			}
		}
		// rounds over all shapes, keeping each shape's fastest run: a clock that is still settling
		// weighs on every shape's early runs instead of on the shapes timed first
		for(int round = 0; round < 3; round++){
			for(Shape &sh : shapes){
				c->chain_cols_set[slot] = sh.C;
				c->chain_strip_set[slot] = (int)sh.M;
				double ms;
				EPHMC_TRY(time_ms(c, e0, e1, 2, &ms, run));
				if(ms < sh.ms) sh.ms = ms;
			}
		}
		ms_out = 1e30;
		for(const Shape &sh : shapes)
			if(sh.ms < ms_out){ ms_out = sh.ms; cols_out = sh.C; strip_out = sh.M; }
		c->chain_cols_set[slot] = cols_out;
		c->chain_strip_set[slot] = (int)strip_out;
		return EPHMC_OK;
	};
	do {
		if((rc = set_zero(c, c->d_xnew, c->dim)) != EPHMC_OK) break;
		// a few untimed chains first, so the first shapes are not timed while the clock rises
		for(int w = 0; w < 4 && rc == EPHMC_OK; w++) rc = md_chain();
		if(rc != EPHMC_OK) break;
		if((rc = tune_chain(prec, md_chain, r.chain_cols, r.chain_strip, r.chain_ms)) != EPHMC_OK) break;
		if(prec != EPHMC_FP64){
			if((rc = tune_chain(EPHMC_FP64, ld_chain, r.chain64_cols, r.chain64_strip, r.chain64_ms)) != EPHMC_OK) break;
		}else{
			r.chain64_cols = r.chain_cols;
			r.chain64_strip = r.chain_strip;
			r.chain64_ms = r.chain_ms;
		}
		// the force sweep below runs on the chain's checkpoints and M_hat^{-1} = 1
		if((rc = md_chain()) != EPHMC_OK) break;
		if((rc = set_zero(c, c->d_mhat, c->mat)) != EPHMC_OK || (rc = add_identity(c, c->d_mhat)) != EPHMC_OK) break;
		// force sweep widths: (4 C + 3) N values of LDS plus the reduction scratch; both orders of
		// the first W phase (its own barrier, or W_0 written ahead)
		struct Sweep { unsigned C; int early; double ms; };
		std::vector<Sweep> sweeps;
		unsigned max_nwg = 0;
		for(unsigned C : {1u, 2u, 4u}){
			if(adjoint_lds(ns, prec, C) > 160 * 1024 || C > ns) continue;
			max_nwg = std::max(max_nwg, (ns + C - 1) / C);
			for(int early = 0; early < 2; early++) sweeps.push_back({C, early, 1e30});
		}
		if(sweeps.empty()){ set_error("autotune: no force-sweep width fits %u sites", ns); rc = EPHMC_ERR_ARG; break; }
		// partials for the widest grid first, then rounds over all shapes as for the chain
		if((rc = reserve_partials(c, max_nwg)) != EPHMC_OK) break;
		for(int round = 0; round < 3 && rc == EPHMC_OK; round++){
			for(Sweep &sw : sweeps){
				c->cols = sw.C;
				c->nwg = (ns + sw.C - 1) / sw.C;
				c->adj_w0_early = sw.early;
				double ms;
				if((rc = time_ms(c, e0, e1, 2, &ms, [&]{ return force_dispatch(c, prec, c->d_mhat, c->d_pdot, nullptr); })) != EPHMC_OK) break;
				if(ms < sw.ms) sw.ms = ms;
			}
		}
		if(rc != EPHMC_OK) break;
		r.adjoint_ms = 1e30;
		int best_early = 1;
		for(const Sweep &sw : sweeps)
			if(sw.ms < r.adjoint_ms){ r.adjoint_ms = sw.ms; r.adjoint_cols = sw.C; best_early = sw.early; }
		c->cols = r.adjoint_cols;
		c->nwg = (ns + c->cols - 1) / c->cols;
		c->adj_w0_early = best_early;
		r.adjoint_w0_early = best_early;
	} while(0);
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	c->timing = timing;
	if(rc != EPHMC_OK){
		c->cols = cols0;
		c->nwg = nwg0;
		c->adj_w0_early = early0;
		for(int q = 0; q < 2; q++){
			c->chain_cols_set[q] = ccols0[q];
			c->chain_strip_set[q] = cstrip0[q];
		}
		return rc;
	}
	if(res) *res = r;
	return EPHMC_OK;
}

int ephmc_timing(ephmc_ctx *c, int enable)
{
	if(!c){ set_error("null context"); return EPHMC_ERR_ARG; }
	EPHMC_HIP(hipStreamSynchronize(c->stream));
	for(auto &kv : c->timers) for(hipEvent_t e : kv.second.ev) (void)hipEventDestroy(e);
	c->timers.clear();
	c->timing = enable != 0;
	return EPHMC_OK;
}

int ephmc_timing_get(ephmc_ctx *c, const char *group, double *avg_ms, unsigned long *calls)
{
	if(!c || !group){ set_error("null argument"); return EPHMC_ERR_ARG; }
	auto it = c->timers.find(group);
	if(it == c->timers.end()){ if(avg_ms) *avg_ms = 0; if(calls) *calls = 0; return EPHMC_OK; }
	TimerGroup &g = it->second;
	EPHMC_TRY(timer_harvest(g));
	if(avg_ms) *avg_ms = g.calls ? g.total_ms / g.calls : 0;
	if(calls) *calls = g.calls;
	return EPHMC_OK;
}

} // extern "C"
