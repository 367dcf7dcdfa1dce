// adjoint.hip -- fermion force as a reverse (adjoint) sweep over the slice product.
//
// The reference force (organic_hmc.c:86-134) needs, for every slice tau, the bond entries
//   Y_tau[i,j] + Y_tau[j,i],   p_dot[tau,i,k] = -kappa_k * flavors * (Y[i,nb] + Y[nb,i]),
// of the exact derivative Y_tau of the slice polynomial p(A_tau) contracted with
//   G_tau = B_{tau+1} ... B_{nt-1} M_hat^{-1} B_0 ... B_{tau-1}
// (construct_all_QM1P + refine_QM1P, organic_dense_matrix.c:311-442).
// With p(A) = sum_n c_n A^n, W_tau = B_{tau+1}..B_{nt-1} M_hat^{-1} and T_tau = B_{tau-1}..B_0
// (so G = W T^T), the derivative is
//   Y = sum_{a+b<=deg-1} c_{a+b+1} (A^a W)(A^b T)^T = sum_a W_a S_a^T,  S_a = sum_b c_{a+b+1} T_b,
// W_a = A^a W, T_b = A^b T.  Bond (i,j): sum_c sum_a (W_a[i,c] S_a[j,c] + W_a[j,c] S_a[i,c]).
// Columns c are independent, so a workgroup owns C of them and produces partial bond sums;
// force_reduce_kernel adds the partials of all workgroups in a fixed order (deterministic
// force, which keeps the molecular dynamics exactly reversible).
//
// Sweep: tau = nt-1 ... 0 with W_{nt-1} = M_hat^{-1}, W_{tau-1} = B_tau W_tau = sum_n c_n W_n.
// T_tau comes from checkpoints T_{s k} stored by the forward chain every k slices and is
// recomputed inside the segment (stored to a workgroup-private scratch, then read backwards).
#include "internal.h"

namespace ephmc {

template <typename R>
struct AdjArgs {
	const R *coef;          // [nt][6][ns]
	const int *nb;          // [6][ns]
	R mu;
	unsigned ns;
	int deg;
	R cm[MAX_DEG + 1];      // monomial coefficients of p
	int np;                 // pair form, for the recompute of T (same arithmetic as the forward chain)
	R a_lin[MAX_PAIRS], a_quad[MAX_PAIRS];
	int cplx[MAX_PAIRS];
	const R *ck;            // checkpoint T_{tau_b} of this segment (column-major N x N); with
	                        // direct, T_tau for every tau of the segment at ck + (tau - tau_b) N^2
	int direct;             // every slice stored: no recompute
	int w0_early;           // W_0 written at the end of the previous slice (no barrier of its own)
	R *seg;                 // scratch: blockIdx.x * (k * C * ns) + (tau - tau_b) * C * ns + c * ns + site
	unsigned seg_stride;    // k * C * ns
	const double *w_init;   // M_hat^{-1} (first segment only), else null
	R *wbuf;                // W state between segments (column-major N x N)
	R *partial;             // [tau - tau_b][nwg][3 ns], in the sweep's precision
	unsigned nwg;
	int tau_b, tau_e;
	// reduction of the previous segment's partials, folded into this launch
	const R *red_src;       // partials of the segment reduced here, or null
	int red_tau_b, red_ntau;
	double f[3];            // -kappa_k flavors
	const double *fscale;   // device scalar multiplying the force (SINGLE_PARTICLE: exp(-logdet)), or null
	double *pdot, *p;       // outputs of the reduction
	double kick;            // p += kick F when p != null
	int clip;               // SCHUR_FORM_INV: bond projections clipped to [-2, 2] (organic_hmc.c:113-114)
};

// Deterministic sum over workgroups of the partial bond sums of one segment:
// F[tau][link] = f_k sum_w partial[tau - tau_b][w][link].  The workgroups are split into
// RED_CHUNKS fixed ranges; every link is the sum of its chunk sums taken in chunk order, so the
// result does not depend on how many threads run the reduction.  The links [l0, l0 + nl) are
// processed in groups of G = nthreads / RED_CHUNKS (at least 1); `scratch` holds
// max(nthreads, RED_CHUNKS) doubles.
constexpr unsigned RED_CHUNKS = 4;

template <typename P>
__device__ void reduce_links(const P *__restrict__ part, unsigned nwg, unsigned loc, unsigned l0, unsigned nl,
		unsigned nlinks_total, int tau_b, const double (&f)[3], const double *__restrict__ fscale, double *__restrict__ pdot,
		double *__restrict__ p, double kick, double *scratch, unsigned nthreads, int clip)
{
	const double fs = fscale ? *fscale : 1.0;
	const unsigned t = threadIdx.x;
	const unsigned CH = nwg < RED_CHUNKS ? nwg : RED_CHUNKS;
	const unsigned G = nthreads / CH ? nthreads / CH : 1;
	for(unsigned g0 = 0; g0 < nl; g0 += G){
		for(unsigned it = t; it < G * CH; it += nthreads){
			const unsigned ll = it % G, ch = it / G;
			double s = 0;
			const unsigned lk = l0 + g0 + ll;
			if(g0 + ll < nl && lk < nlinks_total){
				const unsigned tl = lk / loc, link = lk - tl * loc;
				const unsigned w0 = (unsigned)((size_t)nwg * ch / CH), w1 = (unsigned)((size_t)nwg * (ch + 1) / CH);
				const P *src = part + (size_t)tl * nwg * loc + link;
				// 16 loads in flight, then added in w order (the same sum as one at a time)
				unsigned w = w0;
				for(; w + 16 <= w1; w += 16){
					P v[16];
#pragma unroll
					for(int u = 0; u < 16; u++) v[u] = src[(size_t)(w + u) * loc];
#pragma unroll
					for(int u = 0; u < 16; u++) s += (double)v[u];
				}
				for(; w < w1; w++) s += (double)src[(size_t)w * loc];
			}
			scratch[ch * G + ll] = s;
		}
		__syncthreads();
		for(unsigned ll = t; ll < G; ll += nthreads){
			const unsigned lk2 = l0 + g0 + ll;
			if(g0 + ll < nl && lk2 < nlinks_total){
				double tot = 0;
				for(unsigned q = 0; q < CH; q++) tot += scratch[q * G + ll];
				const unsigned tl = lk2 / loc, link = lk2 - tl * loc;
				if(clip && fabs(tot) > 2.0) tot = tot > 0 ? 2.0 : -2.0;     // NaN stays NaN
				const double F = f[link % 3] * tot * fs;
				const size_t pos = (size_t)(tau_b + tl) * loc + link;
				if(pdot) pdot[pos] = F;
				if(p) p[pos] += kick * F;
			}
		}
		__syncthreads();
	}
}

// A site's C values fit one 16-byte LDS read (fp32 up to 4 columns, fp64 up to 2): with one site
// per thread the register footprint allows the lean form of the sweep (all lanes active, six
// neighbour reads issued together, constant LDS offsets)
template <typename R, int C>
constexpr bool adj_lean(){ return C * sizeof(R) <= 16; }

// BATCH: the six neighbour reads are issued before any of them is used; otherwise each is read
// just before its sums, which holds fewer registers. The sum is taken in the order d = 0..5
// either way.
template <typename R, int C, bool BATCH>
__device__ __forceinline__ void gather6(const R *__restrict__ buf, const int (&nbr)[6], const R (&cf)[6], R mu,
		const R (&own)[C], R (&out)[C])
{
#pragma unroll
	for(int c = 0; c < C; c++) out[c] = mu * own[c];
	if constexpr(BATCH){
		R nv[6][C];
#pragma unroll
		for(int d = 0; d < 6; d++)
#pragma unroll
			for(int c = 0; c < C; c++) nv[d][c] = buf[(size_t)nbr[d] * C + c];
#pragma unroll
		for(int d = 0; d < 6; d++)
#pragma unroll
			for(int c = 0; c < C; c++) out[c] = fmac(cf[d], nv[d][c], out[c]);
	}else{
#pragma unroll
		for(int d = 0; d < 6; d++){
			const R *q = buf + (size_t)nbr[d] * C;
#pragma unroll
			for(int c = 0; c < C; c++) out[c] = fmac(cf[d], q[c], out[c]);
		}
	}
}

template <typename R, int C, int BT, int S>
__host__ __device__ constexpr size_t adj_buf_stride(unsigned ns)
{
	return (sizeof(R) == 4 || (adj_lean<R, C>() && S == 1)) ? (size_t)C * BT * S : (size_t)C * ns;
}

template <typename R, int C, int BT, int S>
__host__ __device__ constexpr size_t adj_lds_bytes(unsigned ns)
{
	return (((4 * adj_buf_stride<R, C, BT, S>(ns) + 3 * (size_t)ns) * sizeof(R) + 7) & ~(size_t)7) + BT * sizeof(double);
}

// DEG is the polynomial degree (a.deg), a template constant so that in the fp32 one-site form the
// phase structure and the S_a recursion need no run-time tests
template <typename R, int C, int BT, int S, int DEG>
__device__ __forceinline__ void adjoint_body(const AdjArgs<R> &a, unsigned bx, unsigned nthr)
{
	extern __shared__ __align__(16) unsigned char smem_raw[];
	const unsigned ns = a.ns;
	// lean one-site form (fp64 only up to degree 4; beyond it, with two sites per thread, or in
	// fp64 at 4 columns, the guarded form and the run-time degree hold fewer registers); fp32
	// always reads the six neighbours together
	constexpr bool LEAN = adj_lean<R, C>() && S == 1 && (sizeof(R) == 4 || DEG <= 4);
	constexpr bool BATCH = sizeof(R) == 4 || LEAN;
	// LDS: buf0, buf1, wb0, wb1 of bs values each, then hx [site][3] and the reduction scratch; in
	// fp32 and in the lean fp64 form bs is the constant C * BT * S, so the four buffers sit at
	// fixed offsets from one neighbour address and a read needs no address add (the offsets fit
	// the 16-bit immediate when C * BT * S * sizeof(R) <= 16 KiB)
	const size_t bs = adj_buf_stride<R, C, BT, S>(ns);
	R *buf0 = reinterpret_cast<R *>(smem_raw);
	R *buf1 = buf0 + bs;
	R *wb0 = buf0 + 2 * bs;
	R *wb1 = buf0 + 3 * bs;
	R *hx = buf0 + 4 * bs;                 // [site][3]
	double *rscratch = reinterpret_cast<double *>(smem_raw + adj_lds_bytes<R, C, BT, S>(ns) - BT * sizeof(double));
	if(a.red_src){
		const unsigned total = (unsigned)a.red_ntau * 3 * ns;
		const unsigned per = (total + a.nwg - 1) / a.nwg;
		reduce_links(a.red_src, a.nwg, 3 * ns, bx * per, per, total, a.red_tau_b, a.f, a.fscale, a.pdot, a.p, a.kick,
			rscratch, nthr, a.clip);
	}
	R *wbs[2] = {wb0, wb1};

	const unsigned col0 = bx * C;
	R *seg = a.seg + (size_t)bx * a.seg_stride;
	unsigned site[S];
	bool act[S];
	int nbr[S][6];
	R cf[S][6];
	R v[S][C];
#pragma unroll
	for(int j = 0; j < S; j++){
		site[j] = threadIdx.x + j * BT;
		act[j] = site[j] < ns;
#pragma unroll
		for(int d = 0; d < 6; d++) nbr[j][d] = act[j] ? a.nb[d * ns + site[j]] : 0;
	}
	auto colok = [&](int c){ return col0 + c < ns; };

	// ---- recompute T_tau, tau in [tau_b, tau_e), from the checkpoint --------------------
	if(!a.direct){
#pragma unroll
		for(int j = 0; j < S; j++)
#pragma unroll
			for(int c = 0; c < C; c++)
				v[j][c] = (act[j] && colok(c)) ? a.ck[(size_t)(col0 + c) * ns + site[j]] : (R)0;
		for(int tau = a.tau_b; tau < a.tau_e; tau++){
			R *dst = seg + (size_t)(tau - a.tau_b) * C * ns;
#pragma unroll
			for(int j = 0; j < S; j++)
				if(act[j])
#pragma unroll
					for(int c = 0; c < C; c++) dst[(size_t)c * ns + site[j]] = v[j][c];
			if(tau == a.tau_e - 1) break;
			const R *cft = a.coef + (size_t)tau * 6 * ns;
#pragma unroll
			for(int j = 0; j < S; j++)
#pragma unroll
				for(int d = 0; d < 6; d++) cf[j][d] = act[j] ? cft[d * ns + site[j]] : (R)0;
			for(int p = 0; p < a.np; p++){
				R w[S][C];
#pragma unroll
				for(int j = 0; j < S; j++)
					if(act[j])
#pragma unroll
						for(int c = 0; c < C; c++) buf0[(size_t)site[j] * C + c] = v[j][c];
				__syncthreads();
#pragma unroll
				for(int j = 0; j < S; j++)
					if(act[j]) gather6<R, C, BATCH>(buf0, nbr[j], cf[j], a.mu, v[j], w[j]);
				if(a.cplx[p]){
#pragma unroll
					for(int j = 0; j < S; j++)
						if(act[j])
#pragma unroll
							for(int c = 0; c < C; c++) buf1[(size_t)site[j] * C + c] = w[j][c];
					__syncthreads();
#pragma unroll
					for(int j = 0; j < S; j++){
						if(!act[j]) continue;
						R z[C];
						gather6<R, C, BATCH>(buf1, nbr[j], cf[j], a.mu, w[j], z);
#pragma unroll
						for(int c = 0; c < C; c++) v[j][c] = v[j][c] + fmac(a.a_quad[p], z[c], a.a_lin[p] * w[j][c]);
					}
				}else{
					__syncthreads();
#pragma unroll
					for(int j = 0; j < S; j++)
#pragma unroll
						for(int c = 0; c < C; c++) v[j][c] = fmac(a.a_lin[p], w[j][c], v[j][c]);
				}
			}
		}

	}

	// ---- W state --------------------------------------------------------------------------
	R wv[S][C];
#pragma unroll
	for(int j = 0; j < S; j++)
#pragma unroll
		for(int c = 0; c < C; c++){
			R val = 0;
			if(act[j] && colok(c)){
				const size_t e = (size_t)(col0 + c) * ns + site[j];
				val = a.w_init ? (R)a.w_init[e] : a.wbuf[e];
			}
			wv[j][c] = val;
		}

	// lean form: every lane loads and gathers, so no lane needs zero values of its own, and the
	// degree is the constant DEG
	constexpr bool ALL_LANES = LEAN;
	const int deg = ALL_LANES ? DEG : a.deg;
	// coefficients and T of slice tau-1 are loaded while slice tau is processed (each thread
	// reads back only the scratch entries it wrote itself)
	R cfn[S][6], tn0[S][C];
	auto load_slice = [&](int tau){
		const R *cft = a.coef + (size_t)tau * 6 * ns;
		const R *tsrc = a.direct ? a.ck + (size_t)(tau - a.tau_b) * ns * ns + (size_t)col0 * ns
		                         : seg + (size_t)(tau - a.tau_b) * C * ns;
#pragma unroll
		for(int j = 0; j < S; j++){
			if constexpr(ALL_LANES){
				// A lane past the last site reads site ns-1 and stores nothing. A column past the
				// last (last workgroup only) reads column ns-1; its W column is zero throughout the
				// sweep, so its products add zeros to the bond sums.
				const unsigned sl = act[j] ? site[j] : ns - 1;
#pragma unroll
				for(int d = 0; d < 6; d++) cfn[j][d] = cft[d * ns + sl];
#pragma unroll
				for(int c = 0; c < C; c++) tn0[j][c] = tsrc[(size_t)(colok(c) ? c : ns - 1 - col0) * ns + sl];
			}else{
#pragma unroll
				for(int d = 0; d < 6; d++) cfn[j][d] = act[j] ? cft[d * ns + site[j]] : (R)0;
#pragma unroll
				for(int c = 0; c < C; c++) tn0[j][c] = (act[j] && colok(c)) ? tsrc[(size_t)c * ns + site[j]] : (R)0;
			}
		}
	};
	// bond(i, 2k) = H[i][2k] + H[nb(i,2k)][2k+1]: the H[.][2k+1] of a slice go to LDS (hx) at
	// its end and, in FP32, the bond sums are written after the next slice's first barrier
	// (after the last slice, after one more barrier); FP64 keeps a barrier of its own per slice,
	// as holding H[.][2k] across the slice would spill registers
	constexpr bool FOLD = sizeof(R) == 4;
	R hb[S][3];
	int tau_out = -1;
	auto flush_bonds = [&](){
		R *out = a.partial + ((size_t)(tau_out - a.tau_b) * a.nwg + bx) * 3 * ns;
#pragma unroll
		for(int j = 0; j < S; j++)
			if(act[j])
#pragma unroll
				for(int k = 0; k < 3; k++) out[(size_t)site[j] * 3 + k] = hb[j][k] + hx[(size_t)nbr[j][2 * k] * 3 + k];
	};
	// W_0 of a slice goes to wbs[0] as soon as it is known (here, then at the end of each slice);
	// the T-power barriers make it visible, so the first W phase needs no barrier of its own
	auto put_w0 = [&](){
#pragma unroll
		for(int j = 0; j < S; j++)
			if(act[j])
#pragma unroll
				for(int c = 0; c < C; c++) wbs[0][(size_t)site[j] * C + c] = wv[j][c];
	};
	if(a.w0_early) put_w0();
	load_slice(a.tau_e - 1);
	for(int tau = a.tau_e - 1; tau >= a.tau_b; tau--){
		R t[S][C];
#pragma unroll
		for(int j = 0; j < S; j++){
#pragma unroll
			for(int d = 0; d < 6; d++) cf[j][d] = cfn[j][d];
#pragma unroll
			for(int c = 0; c < C; c++) t[j][c] = tn0[j][c];
		}
		if(tau > a.tau_b) load_slice(tau - 1);

		// S_a = sum_b c_{a+b+1} T_b, built while the powers T_b = A^b T are formed
		R sa[DEG][S][C];
#pragma unroll
		for(int q = 0; q < DEG; q++)
#pragma unroll
			for(int j = 0; j < S; j++)
#pragma unroll
				for(int c = 0; c < C; c++) sa[q][j][c] = (q < deg) ? a.cm[q + 1] * t[j][c] : (R)0;
		// unrolled: b is a constant, so bb is a known LDS buffer
#pragma unroll
		for(int b = 1; b < DEG; b++){
			if(b >= deg) break;
			R *bb = ((b - 1) & 1) ? buf1 : buf0;
#pragma unroll
			for(int j = 0; j < S; j++)
				if(act[j])
#pragma unroll
					for(int c = 0; c < C; c++) bb[(size_t)site[j] * C + c] = t[j][c];
			__syncthreads();
			if(FOLD && a.w0_early && b == 1 && tau_out >= 0) flush_bonds();
#pragma unroll
			for(int j = 0; j < S; j++){
				// with ALL_LANES a thread past the last site gathers too (neighbour 0, in bounds)
				// and stores nothing
				R tn[C];
				if(ALL_LANES || act[j]) gather6<R, C, BATCH>(bb, nbr[j], cf[j], a.mu, t[j], tn);
				else
#pragma unroll
					for(int c = 0; c < C; c++) tn[c] = 0;
#pragma unroll
				for(int c = 0; c < C; c++) t[j][c] = tn[c];
			}
#pragma unroll
			for(int q = 0; q < DEG; q++)
				if(q + b < deg)
#pragma unroll
					for(int j = 0; j < S; j++)
#pragma unroll
						for(int c = 0; c < C; c++) sa[q][j][c] = fmac(a.cm[q + b + 1], t[j][c], sa[q][j][c]);
		}

		// W powers with the half-bond sums taken from the same neighbour reads:
		// H[i][d] = sum_{q,c} S_q[i,c] W_q[nb(i,d),c]
		R h[S][6];
		R acc[S][C], wa[S][C];
#pragma unroll
		for(int j = 0; j < S; j++){
#pragma unroll
			for(int d = 0; d < 6; d++) h[j][d] = 0;
#pragma unroll
			for(int c = 0; c < C; c++){
				acc[j][c] = a.cm[0] * wv[j][c];
				wa[j][c] = wv[j][c];
			}
		}
#pragma unroll
		for(int q = 0; q < DEG; q++){
			if(q >= deg) break;
			R *wb = wbs[q & 1];
			if(q > 0 || !a.w0_early){
#pragma unroll
				for(int j = 0; j < S; j++)
					if(act[j])
#pragma unroll
						for(int c = 0; c < C; c++) wb[(size_t)site[j] * C + c] = wa[j][c];
				__syncthreads();
				if(FOLD && !a.w0_early && q == 0 && tau_out >= 0) flush_bonds();
			}
#pragma unroll
			for(int j = 0; j < S; j++){
				if(!act[j]) continue;
				R wn[C];
				// BATCH: all six neighbour reads are issued before any of them is used (otherwise
				// each is read just before its sums; same arithmetic)
				constexpr int NV = BATCH ? 6 : 1;
				R nv[NV][C];
				if constexpr(NV == 6){
#pragma unroll
					for(int d = 0; d < 6; d++)
#pragma unroll
						for(int c = 0; c < C; c++) nv[d][c] = wb[(size_t)nbr[j][d] * C + c];
				}
#pragma unroll
				for(int c = 0; c < C; c++) wn[c] = a.mu * wa[j][c];
#pragma unroll
				for(int d = 0; d < 6; d++){
					const int dv = NV == 6 ? d : 0;
					if constexpr(NV == 1){
#pragma unroll
						for(int c = 0; c < C; c++) nv[0][c] = wb[(size_t)nbr[j][d] * C + c];
					}
					R hs = 0;
#pragma unroll
					for(int c = 0; c < C; c++){
						hs = fmac(sa[q][j][c], nv[dv][c], hs);
						wn[c] = fmac(cf[j][d], nv[dv][c], wn[c]);
					}
					// (with ALL_LANES the first phase's sum is taken as is; the guarded form adds
					// it to zero, which holds fewer registers)
					h[j][d] = (ALL_LANES && q == 0) ? hs : h[j][d] + hs;
				}
#pragma unroll
				for(int c = 0; c < C; c++){
					wa[j][c] = wn[c];
					acc[j][c] = fmac(a.cm[q + 1], wn[c], acc[j][c]);
				}
			}
		}
#pragma unroll
		for(int j = 0; j < S; j++){
			if(act[j])
#pragma unroll
				for(int k = 0; k < 3; k++) hx[(size_t)site[j] * 3 + k] = h[j][2 * k + 1];
#pragma unroll
			for(int k = 0; k < 3; k++) hb[j][k] = h[j][2 * k];
		}
		tau_out = tau;
		if(!FOLD){
			__syncthreads();
			flush_bonds();
			tau_out = -1;
		}
#pragma unroll
		for(int j = 0; j < S; j++)
#pragma unroll
			for(int c = 0; c < C; c++) wv[j][c] = acc[j][c];
		// wbs[0] was last read in W phase deg - 2 (deg is even), before the barrier of phase deg - 1
		if(a.w0_early){
			if((deg - 1) % 2 == 0) __syncthreads();
			put_w0();
		}
	}
	__syncthreads();
	if(tau_out >= 0) flush_bonds();
#pragma unroll
	for(int j = 0; j < S; j++)
#pragma unroll
		for(int c = 0; c < C; c++)
			if(act[j] && colok(c)) a.wbuf[(size_t)(col0 + c) * ns + site[j]] = wv[j][c];
}

template <typename R, int C, int BT, int S, int DEG>
__global__ __launch_bounds__(BT) void adjoint_kernel(AdjArgs<R> a)
{
	adjoint_body<R, C, BT, S, DEG>(a, blockIdx.x, BT);
}

// many chains in one launch: blockIdx.z = chain (its record in tab), blockIdx.x = its workgroup.
// With one site per thread the workgroup may have fewer threads than BT (at least ns): the LDS
// layout stays that of BT, and the reduction's sums do not depend on the thread count.
template <typename R, int C, int BT, int S, int DEG>
__global__ __launch_bounds__(BT) void adjoint_multi_kernel(const AdjArgs<R> *__restrict__ tab)
{
	static_assert(S == 1, "one site per thread");
	const AdjArgs<R> &a = tab[blockIdx.z];
	if(blockIdx.x >= a.nwg) return;
	adjoint_body<R, C, BT, S, DEG>(a, blockIdx.x, blockDim.x);
}

template <typename P>
__global__ __launch_bounds__(256) void force_reduce_kernel(const P *__restrict__ partial, unsigned nwg, unsigned loc,
		int tau_b, int ntau, double f0, double f1, double f2, const double *__restrict__ fscale, double *__restrict__ pdot,
		double *__restrict__ p, double kick, int clip)
{
	__shared__ double scratch[256];
	const double f[3] = {f0, f1, f2};
	const unsigned total = (unsigned)ntau * loc;
	reduce_links(partial, nwg, loc, blockIdx.x * 64, 64, total, tau_b, f, fscale, pdot, p, kick, scratch, 256, clip);
}

template <typename P>
struct ReduceArgs {
	const P *partial;
	unsigned nwg, loc;
	int tau_b, ntau;
	double f[3];
	const double *fscale;
	double *pdot, *p;
	double kick;
	int clip;
	unsigned gx;
};

template <typename P>
__global__ __launch_bounds__(256) void force_reduce_multi_kernel(const ReduceArgs<P> *__restrict__ tab)
{
	__shared__ double scratch[256];
	const ReduceArgs<P> &a = tab[blockIdx.z];
	if(blockIdx.x >= a.gx) return;
	const unsigned total = (unsigned)a.ntau * a.loc;
	reduce_links(a.partial, a.nwg, a.loc, blockIdx.x * 64, 64, total, a.tau_b, a.f, a.fscale, a.pdot, a.p, a.kick, scratch,
		256, a.clip);
}

// CONSTANT_FORCE: every slice gets the force of slice 0 (organic_hmc.c:92-98,128)
__global__ void broadcast_slice_kernel(double *__restrict__ pdot, unsigned loc, unsigned nt)
{
	const size_t n = (size_t)loc * nt;
	for(size_t e = loc + blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
		pdot[e] = pdot[e % loc];
}

// LDS bytes of the force sweep at C columns per workgroup, for the launch shape that
// launch_adjoint_c picks for ns sites (the same rule, so a width chosen by this size launches)
template <typename R, int C>
static size_t adj_lds_for(unsigned ns)
{
	if(ns <= 256) return adj_lds_bytes<R, C, 256, 1>(ns);
	if(ns <= 512) return adj_lds_bytes<R, C, 512, 1>(ns);
	if(ns <= 1024 && sites_per_thread("EPHMC_ADJOINT_SITES") == 2) return adj_lds_bytes<R, C, 512, 2>(ns);
	if(ns <= 1024) return adj_lds_bytes<R, C, 1024, 1>(ns);
	return adj_lds_bytes<R, C, 1024, 2>(ns);
}

size_t adjoint_lds(unsigned ns, int prec, unsigned C)
{
	if(prec == EPHMC_FP64){
		switch(C){
			case 1: return adj_lds_for<double, 1>(ns);
			case 2: return adj_lds_for<double, 2>(ns);
			default: return adj_lds_for<double, 4>(ns);
		}
	}
	switch(C){
		case 1: return adj_lds_for<float, 1>(ns);
		case 2: return adj_lds_for<float, 2>(ns);
		default: return adj_lds_for<float, 4>(ns);
	}
}

unsigned adjoint_cols(const ephmc_ctx *c, int prec)
{
	return c->cols ? c->cols : (prec == EPHMC_FP64 ? 2u : 4u);
}

template <typename R, int C>
static int launch_adjoint_c(ephmc_ctx *c, AdjArgs<R> a)
{
	const unsigned ns = a.ns;
	const unsigned blocks = a.nwg;
#define EPHMC_ADJ_LAUNCH2(BT, S, DEG)                                                             \
	do {                                                                                          \
		const size_t lds = adj_lds_bytes<R, C, BT, S>(ns);                                        \
		if(lds > 160 * 1024){ set_error("adjoint: %zu B of LDS exceeds 160 KiB", lds); return EPHMC_ERR_ARG; } \
		(void)hipFuncSetAttribute((const void *)adjoint_kernel<R, C, BT, S, DEG>,                       \
			hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                                \
		adjoint_kernel<R, C, BT, S, DEG><<<blocks, BT, lds, c->stream>>>(a);                     \
	} while(0)
#define EPHMC_ADJ_LAUNCH(BT, S)                                                                   \
	do {                                                                                          \
		switch(a.deg){                                                                            \
			case 2: EPHMC_ADJ_LAUNCH2(BT, S, 2); break;                                               \
			case 4: EPHMC_ADJ_LAUNCH2(BT, S, 4); break;                                               \
			case 6: EPHMC_ADJ_LAUNCH2(BT, S, 6); break;                                               \
			case 8: EPHMC_ADJ_LAUNCH2(BT, S, 8); break;                                               \
			default: set_error("adjoint: polynomial degree %d (Taylor orders 2, 4, 6, 8)", a.deg);     \
				return EPHMC_ERR_UNSUPPORTED;                                                         \
		}                                                                                         \
	} while(0)
	if(ns <= 256) EPHMC_ADJ_LAUNCH(256, 1);
	else if(ns <= 512) EPHMC_ADJ_LAUNCH(512, 1);
	else if(ns <= 1024 && sites_per_thread("EPHMC_ADJOINT_SITES") == 2) EPHMC_ADJ_LAUNCH(512, 2);
	else if(ns <= 1024) EPHMC_ADJ_LAUNCH(1024, 1);
	else if(ns <= 2048) EPHMC_ADJ_LAUNCH(1024, 2);
	else { set_error("adjoint: lattice of %u sites exceeds 2048", ns); return EPHMC_ERR_UNSUPPORTED; }
#undef EPHMC_ADJ_LAUNCH
#undef EPHMC_ADJ_LAUNCH2
	EPHMC_LAUNCHED();
	return EPHMC_OK;
}

// Arguments of the force sweep common to all its launches (per-segment fields: adj_segment_args).
template <typename R>
static AdjArgs<R> adj_base_args(ephmc_ctx *c, int prec, unsigned C, double *pdot, double *p, double kick,
		const double *fscale)
{
	const unsigned ns = c->ns;
	AdjArgs<R> a{};
	a.coef = prec == EPHMC_FP64 ? (const R *)c->d_coef64 : (const R *)c->d_coef32;
	a.nb = c->d_nb;
	a.mu = (R)c->cfg.mu;
	a.ns = ns;
	a.deg = c->poly.deg;
	for(int q = 0; q <= c->poly.deg; q++) a.cm[q] = (R)c->poly.c[q];
	a.np = c->poly.np;
	for(int q = 0; q < c->poly.np; q++){
		a.a_lin[q] = (R)c->poly.a_lin[q];
		a.a_quad[q] = (R)c->poly.a_quad[q];
		a.cplx[q] = c->poly.cplx[q];
	}
	a.seg_stride = c->ckpt * C * ns;
	a.nwg = (ns + C - 1) / C;
	// W state between segments; T_nt is no longer needed once M_hat is formed
	a.wbuf = prec == EPHMC_FP64 ? (R *)(void *)c->d_Tfin : (R *)(void *)c->d_Tfin32;
	for(int q = 0; q < 3; q++) a.f[q] = -c->cfg.kappa[q] * c->cfg.flavors;
	a.pdot = pdot;
	a.p = p;
	a.kick = kick;
	a.fscale = fscale;
	a.clip = c->cfg.schur_form_inv != 0;
	return a;
}

// Segment s of nseg (slices [s span, (s+1) span) up to nt): it starts from minv when it is the
// last one and reduces the partials of the segment before it (prev, or none)
template <typename R>
static void adj_segment_args(ephmc_ctx *c, AdjArgs<R> &a, int s, unsigned nseg, const double *minv, const R *prev,
		int prev_b, int prev_n, unsigned nt)
{
	const unsigned k = c->ckpt, span = c->span;
	a.tau_b = s * (int)span;
	a.tau_e = std::min<int>((s + 1) * (int)span, (int)nt);
	a.ck = (const R *)c->d_ck + (size_t)(a.tau_b / (int)k) * c->mat;
	a.direct = k == 1;
	a.w0_early = c->adj_w0_early;
	a.seg = (R *)c->d_seg;
	a.w_init = (s == (int)nseg - 1) ? minv : nullptr;
	a.partial = (R *)(void *)c->d_partial[s & 1];
	a.red_src = prev;
	a.red_tau_b = prev_b;
	a.red_ntau = prev_n;
}

// Segments s = nseg-1 .. 0 of the reverse sweep; W starts from minv (fp64) or, when minv is null,
// from the W buffer (already in the precision R).  nt_eff < nt restricts the sweep to slices
// [0, nt_eff) (CONSTANT_FORCE: nt_eff = 1 with W_0 prepared by the caller).
template <typename R>
static int run_adjoint(ephmc_ctx *c, int prec, const double *minv, double *pdot, double kick, double *p,
		const double *fscale, unsigned nt_eff)
{
	const unsigned nt = nt_eff, span = c->span, C = adjoint_cols(c, prec);
	AdjArgs<R> a = adj_base_args<R>(c, prec, C, pdot, p, kick, fscale);
	const double f[3] = {a.f[0], a.f[1], a.f[2]};
	const unsigned nseg = (nt + span - 1) / span;
	int prev_b = 0, prev_n = 0;
	const R *prev = nullptr;
	for(int s = (int)nseg - 1; s >= 0; s--){
		adj_segment_args<R>(c, a, s, nseg, minv, prev, prev_b, prev_n, nt);
		int rc;
		EPHMC_TRY(timer_begin(c, "adjoint_kernel"));
		switch(C){
			case 1: rc = launch_adjoint_c<R, 1>(c, a); break;
			case 2: rc = launch_adjoint_c<R, 2>(c, a); break;
			default: rc = launch_adjoint_c<R, 4>(c, a); break;
		}
		if(rc != EPHMC_OK) return rc;
		EPHMC_TRY(timer_end(c, "adjoint_kernel"));
		prev = a.partial;
		prev_b = a.tau_b;
		prev_n = a.tau_e - a.tau_b;
	}
	const size_t n = (size_t)prev_n * c->loc;
	EPHMC_TRY(timer_begin(c, "force_reduce"));
	force_reduce_kernel<R><<<(unsigned)((n + 63) / 64), 256, 0, c->stream>>>(prev, a.nwg, c->loc, prev_b, prev_n,
		f[0], f[1], f[2], fscale, pdot, p, kick, a.clip);
	EPHMC_LAUNCHED();
	EPHMC_TRY(timer_end(c, "force_reduce"));
	return EPHMC_OK;
}

// Force for the field whose coefficients and checkpoints are current in the context
// (chain_forward with checkpoints must have run for this field).
int force_adjoint(ephmc_ctx *c, int prec, const double *minv, double *pdot, double kick, double *p, const double *fscale)
{
	if(prec == EPHMC_FP64) return run_adjoint<double>(c, prec, minv, pdot, kick, p, fscale, c->nt);
	return run_adjoint<float>(c, prec, minv, pdot, kick, p, fscale, c->nt);
}

// ---- many chains in one launch (batch.hip): the sweep of force_adjoint(c, prec, minv, pdot, 0,
// null, null) with one site per thread on `block` threads (ns <= block <= threads, the LDS layout's
// workgroup size); the bits do not depend on the workgroup size (the columns per workgroup, which
// fix the partial sums, are the context's)
template <typename R, int C, int BT>
static int batch_adj_add(BatchOp &op, unsigned block, const AdjArgs<R> &a)
{
	const size_t lds = adj_lds_bytes<R, C, BT, 1>(a.ns);
	if(lds > 160 * 1024){ set_error("batch adjoint: %zu B of LDS exceeds 160 KiB", lds); return EPHMC_ERR_ARG; }
	const void *fn = (const void *)adjoint_multi_kernel<R, C, BT, 1, 4>;
	(void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)std::max(lds, op.lds));
	return op.add(fn, block, lds, a.nwg, 1, a);
}

template <typename R, int C>
static int batch_adj_bt(BatchOp &op, unsigned threads, unsigned block, const AdjArgs<R> &a)
{
	switch(threads){
		case 256: return batch_adj_add<R, C, 256>(op, block, a);
		case 512: return batch_adj_add<R, C, 512>(op, block, a);
		case 1024: return batch_adj_add<R, C, 1024>(op, block, a);
		default: set_error("batch adjoint: %u threads per workgroup", threads); return EPHMC_ERR_ARG;
	}
}

template <typename R>
static int batch_adjoint_r(std::vector<BatchOp> &segs, BatchOp &reduce, ephmc_ctx *c, int prec, unsigned threads,
		unsigned block, const double *minv, double *pdot)
{
	if(c->poly.deg != 4){ set_error("batch: polynomial degree %d (the batch runs Taylor order 4)", c->poly.deg); return EPHMC_ERR_UNSUPPORTED; }
	if(c->ns > block || block > threads || block % 64){
		set_error("batch adjoint: %u sites on %u of %u threads", c->ns, block, threads);
		return EPHMC_ERR_ARG;
	}
	const unsigned nt = c->nt, span = c->span, C = adjoint_cols(c, prec);
	AdjArgs<R> a = adj_base_args<R>(c, prec, C, pdot, nullptr, 0, nullptr);
	const unsigned nseg = (nt + span - 1) / span;
	if(segs.size() < nseg) segs.resize(nseg);
	int prev_b = 0, prev_n = 0;
	const R *prev = nullptr;
	for(int s = (int)nseg - 1; s >= 0; s--){
		adj_segment_args<R>(c, a, s, nseg, minv, prev, prev_b, prev_n, nt);
		BatchOp &op = segs[nseg - 1 - s];
		int rc;
		switch(C){
			case 1: rc = batch_adj_bt<R, 1>(op, threads, block, a); break;
			case 2: rc = batch_adj_bt<R, 2>(op, threads, block, a); break;
			default: rc = batch_adj_bt<R, 4>(op, threads, block, a); break;
		}
		if(rc != EPHMC_OK) return rc;
		prev = a.partial;
		prev_b = a.tau_b;
		prev_n = a.tau_e - a.tau_b;
	}
	ReduceArgs<R> r{};
	r.partial = prev;
	r.nwg = a.nwg;
	r.loc = c->loc;
	r.tau_b = prev_b;
	r.ntau = prev_n;
	for(int q = 0; q < 3; q++) r.f[q] = a.f[q];
	r.pdot = pdot;
	r.clip = a.clip;
	r.gx = (unsigned)(((size_t)prev_n * c->loc + 63) / 64);       // the grid of run_adjoint's reduction
	return reduce.add((const void *)force_reduce_multi_kernel<R>, 256, 0, r.gx, 1, r);
}

int batch_adjoint(std::vector<BatchOp> &segs, BatchOp &reduce, ephmc_ctx *c, int prec, unsigned threads,
		unsigned block, const double *minv, double *pdot)
{
	if(prec == EPHMC_FP64) return batch_adjoint_r<double>(segs, reduce, c, prec, threads, block, minv, pdot);
	return batch_adjoint_r<float>(segs, reduce, c, prec, threads, block, minv, pdot);
}

// CONSTANT_FORCE (organic_hmc.c:92-98 with construct_QM1P at time 0, organic_dense_matrix.c:267-309):
// W_0 = B_1 ... B_{nt-1} M_hat^{-1} (or the identity chain for SINGLE_PARTICLE), the exact
// slice-0 derivative with T_0 = 1 (checkpoint 0 of the forward chain), the same force on every
// slice.  minv: M_hat^{-1} (fp64), or null for SINGLE_PARTICLE.
int force_constant(ephmc_ctx *c, int prec, const double *minv, double *pdot, const double *fscale)
{
	void *w0 = prec == EPHMC_FP64 ? (void *)c->d_Tfin : (void *)c->d_Tfin32;
	EPHMC_TRY(chain_range(c, prec, minv, w0, (int)c->nt - 1, -1, (int)c->nt - 1));
	if(prec == EPHMC_FP64) EPHMC_TRY(run_adjoint<double>(c, prec, nullptr, pdot, 0, nullptr, fscale, 1));
	else EPHMC_TRY(run_adjoint<float>(c, prec, nullptr, pdot, 0, nullptr, fscale, 1));
	return broadcast_slice0(c, pdot);
}

// every slice gets the force of slice 0
int broadcast_slice0(ephmc_ctx *c, double *pdot)
{
	if(c->nt > 1){
		broadcast_slice_kernel<<<1024, 256, 0, c->stream>>>(pdot, c->loc, c->nt);
		EPHMC_LAUNCHED();
	}
	return EPHMC_OK;
}

} // namespace ephmc
