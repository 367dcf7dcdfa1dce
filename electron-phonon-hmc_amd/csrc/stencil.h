// stencil.h -- strip form of the slice stencil shared by the chain and the adjoint sweep.
//
// A thread owns M consecutive sites s0 .. s0+M-1 of one lattice row (l1 % M == 0).  Inside the
// strip the left/right neighbours are the thread's own registers, and neighbouring sites share
// their neighbours in the rows above and below:
//   up[k], k = 0..M: TL(s0), TR(s0), TR(s0+1), ..., TR(s0+M-1)   (row above, a0-1 .. a0+M-1)
//   dn[k], k = 0..M: BL(s0), ..., BL(s0+M-1), BR(s0+M-1)         (row below, a0 .. a0+M)
//   lf = L(s0), rt = R(s0+M-1)
// so site k has R = own[k+1] (or rt), L = own[k-1] (or lf), TR = up[k+1], BL = dn[k],
// TL = up[k], BR = dn[k+1]: 2M+4 LDS reads of C values per strip instead of 6M.  The identities
// TL(s0+k) = TR(s0+k-1) and BR(s0+k) = BL(s0+k+1) hold because the strip does not wrap around
// the row.  Buffers are [site][C] in LDS.
#pragma once
#include "internal.h"

namespace ephmc {

template <int M>
struct StripNb {
	int up[M + 1], dn[M + 1], lf, rt;
};

template <int M>
__device__ __forceinline__ void strip_neighbours(const int *__restrict__ nb, unsigned ns, unsigned s0, StripNb<M> &sn)
{
	sn.up[0] = nb[4 * ns + s0];
#pragma unroll
	for(int k = 0; k < M; k++){
		sn.up[k + 1] = nb[2 * ns + s0 + k];
		sn.dn[k] = nb[3 * ns + s0 + k];
	}
	sn.dn[M] = nb[5 * ns + s0 + M - 1];
	sn.lf = nb[1 * ns + s0];
	sn.rt = nb[0 * ns + s0 + M - 1];
}

template <typename R, int C, int M>
__device__ __forceinline__ void strip_gather(const R *__restrict__ buf, const StripNb<M> &sn, const R (&cf)[M][6], R mu,
		const R (&own)[M][C], R (&out)[M][C])
{
	R up[M + 1][C], dn[M + 1][C], lf[C], rt[C];
#pragma unroll
	for(int k = 0; k <= M; k++)
#pragma unroll
		for(int c = 0; c < C; c++){
			up[k][c] = buf[(size_t)sn.up[k] * C + c];
			dn[k][c] = buf[(size_t)sn.dn[k] * C + c];
		}
#pragma unroll
	for(int c = 0; c < C; c++){
		lf[c] = buf[(size_t)sn.lf * C + c];
		rt[c] = buf[(size_t)sn.rt * C + c];
	}
#pragma unroll
	for(int k = 0; k < M; k++)
#pragma unroll
		for(int c = 0; c < C; c++){
			const R r = k < M - 1 ? own[k + 1][c] : rt[c];
			const R l = k > 0 ? own[k - 1][c] : lf[c];
			R o = mu * own[k][c];
			o = fmac(cf[k][0], r, o);
			o = fmac(cf[k][1], l, o);
			o = fmac(cf[k][2], up[k + 1][c], o);
			o = fmac(cf[k][3], dn[k][c], o);
			o = fmac(cf[k][4], up[k][c], o);
			o = fmac(cf[k][5], dn[k + 1][c], o);
			out[k][c] = o;
		}
}


// The six neighbour values of every strip site (d = R, L, TR, BL, TL, BR) from buf and the
// thread's own values, for contractions other than the stencil (the half-bond sums).
template <typename R, int C, int M>
struct StripVals {
	R up[M + 1][C], dn[M + 1][C], lf[C], rt[C];
};

template <typename R, int C, int M>
__device__ __forceinline__ void strip_load(const R *__restrict__ buf, const StripNb<M> &sn, StripVals<R, C, M> &sv)
{
#pragma unroll
	for(int k = 0; k <= M; k++)
#pragma unroll
		for(int c = 0; c < C; c++){
			sv.up[k][c] = buf[(size_t)sn.up[k] * C + c];
			sv.dn[k][c] = buf[(size_t)sn.dn[k] * C + c];
		}
#pragma unroll
	for(int c = 0; c < C; c++){
		sv.lf[c] = buf[(size_t)sn.lf * C + c];
		sv.rt[c] = buf[(size_t)sn.rt * C + c];
	}
}

} // namespace ephmc
