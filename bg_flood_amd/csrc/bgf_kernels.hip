// bgf_kernels.hip -- HIP/CDNA4 kernels of the BG_Flood Kurganov timestep on
// a one-level block-uniform grid.
//
// One workgroup of 256 threads (four wave64s) owns one 16x16 block; thread t
// owns cell (t % 16, t / 16).  Per step:
//
//   k_predictor  gathers h, u, v, zb on a 20x20 tile (two cells of each
//                neighbour block, or the block's own edge cell at a domain
//                boundary) straight from the neighbours into LDS: this is the
//                halo fill of src/Halo.cu:586-616 plus the distance-2 reads of
//                gradientHalo (src/Gradients.cu:2713-3294) without a halo
//                pass.  It then forms the theta-minmod gradients
//                (src/Gradients.cu:921-957), the 17x16 X and 16x17 Y
//                Kurganov-Petrova faces (src/Kurganov.cu:474-641,
//                1404-1546), the wall patch (src/Boundary.cu:587-760), the
//                tendencies (src/Advection.cu:160-254), and the block's
//                minimum CFL step (src/Advection.cu:470-498).
//   k_reduce_dt  one workgroup: global minimum, 1.5x growth clamp and
//                output-time landing in FP64 (src/Advection.cu:511-535),
//                then totaltime += dt (src/Mainloop.cu:60) on the GPU.
//   k_corrector  rebuilds U* = U + dt/2 L(U) (src/Advection.cu:332-385) on
//                its own 20x20 tile from U and the predictor tendencies, so
//                no U* array is written; forms the corrector fluxes of U*,
//                advances U^{n+1} = U + dt L(U*), applies bottom friction
//                (src/Friction.cu:89-148) and implicit rain
//                (src/Updateforcing.cu:530-581), and writes h, u, v to the
//                other buffer of a ping-pong pair.
//
// HBM traffic per cell-update: predictor reads h, u, v, zb and writes
// dh, dhu, dhv; corrector reads h, u, v, dh, dhu, dhv, zb, cf and writes
// h, u, v: 18 words.  Neighbour strips are re-read from cache.
//
// EXACT kernels (FAST = false) perform every operation of the reference CPU
// step in the same order and precision (this file is compiled with
// -ffp-contract=off; divisions and square roots are IEEE), so they reproduce
// the CPU step bit for bit.  FAST kernels use the hardware reciprocal and
// square root and fused multiply-adds.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bgf_kernels.h"

namespace bgf {

constexpr int BW = 16;
constexpr int MW = 18;
constexpr int BS = MW * MW;
constexpr int TP = 20;  // LDS tile pitch: x, y in [-2, 17]

__device__ __forceinline__ int memloc(int i, int j, int ib) { return (i + 1) + (j + 1) * MW + ib * BS; }
__device__ __forceinline__ int tix(int x, int y) { return (y + 2) * TP + (x + 2); }
__device__ __forceinline__ int gxi(int x, int y) { return y * 18 + (x + 1); }   // x in [-1,16], y in [0,15]
__device__ __forceinline__ int gyi(int x, int y) { return (y + 1) * 16 + x; }   // x in [0,15], y in [-1,16]
__device__ __forceinline__ int fxi(int xf, int y) { return y * 17 + xf; }       // xf in [0,16], y in [0,15]
__device__ __forceinline__ int fyi(int x, int yf) { return yf * 16 + x; }       // x in [0,15], yf in [0,16]

template <class T> __device__ __forceinline__ T dmax(T a, T b) { return a > b ? a : b; }
template <class T> __device__ __forceinline__ T dmin(T a, T b) { return a < b ? a : b; }

// Division and square root: IEEE in EXACT kernels, hardware approximations in FAST ones.
template <class T, bool FAST> struct Ar {
    __device__ static __forceinline__ T div(T a, T b) { return a / b; }
    __device__ static __forceinline__ T rcp(T a) { return T(1.0) / a; }
    __device__ static __forceinline__ T sqrt_(T a) { return sqrt(a); }
};
template <> struct Ar<float, true> {
    __device__ static __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    __device__ static __forceinline__ float rcp(float a) { return __builtin_amdgcn_rcpf(a); }
    __device__ static __forceinline__ float sqrt_(float a) { return __builtin_amdgcn_sqrtf(a); }
};

// src/Util_CPU.cu:243-263
template <class T> __device__ __forceinline__ T minmod2(T theta, T s0, T s1, T s2) {
    if (s0 < s1 && s1 < s2) {
        T d1 = theta * (s1 - s0);
        T d2 = (s2 - s0) / T(2.0);
        T d3 = theta * (s2 - s1);
        if (d2 < d1) d1 = d2;
        return dmin(d1, d3);
    }
    if (s0 > s1 && s1 > s2) {
        T d1 = theta * (s1 - s0), d2 = (s2 - s0) / T(2.0), d3 = theta * (s2 - s1);
        if (d2 > d1) d1 = d2;
        return dmax(d1, d3);
    }
    return T(0.0);
}

template <class T> struct Cst {
    T g, ga, CFL, theta, epsi, inv_epsi, eps_w, eps, delta, dxh, cmdinv, fc, rainrate;
    double eps_d;
};

template <class T, bool FAST> __device__ __forceinline__ Cst<T> make_cst(const KParams& P) {
    Cst<T> c;
    c.g = T(P.g);
    c.ga = c.g * T(0.5);
    c.CFL = T(P.CFL);
    c.theta = T(P.theta);
    c.epsi = sizeof(T) == 8 ? T(2.220446049250313e-16) : T(1.1920928955078125e-07);
    c.inv_epsi = T(1.0) / c.epsi;
    c.eps_w = T(P.eps) + c.epsi;
    c.eps = T(P.eps);
    c.eps_d = P.eps;
    c.delta = T(P.delta);
    c.dxh = c.delta * T(0.5);
    c.cmdinv = T(1.0) / (T(1.0) * c.delta);
    c.fc = T(P.fc);
    // rain depth rate; the per-step depth is max(R/1000/3600*dt, 0)
    c.rainrate = T(P.rain_mmh);
    return c;
}

// Hydrostatic reconstruction and central-upwind flux at one face between the
// "n" (left/bottom) cell and the "i" (right/top) cell.  Along Y the roles of
// u and v swap.  src/Kurganov.cu:62-164 and 1865-1911.
template <class T, bool FAST>
__device__ __forceinline__ bool kurg_face(const Cst<T>& C, T hi, T hn, T zsi, T zsn, T dhi, T dhn, T dzsi, T dzsn,
                                          T ui, T un, T dui, T dun, T vi, T vn, T dvi, T dvn,
                                          T& Fh, T& Fq, T& S, T& Fv, T& dt) {
    typedef Ar<T, FAST> A;
    if (hi > C.eps_w || hn > C.eps_w) {
        const T dx = C.dxh;
        T zi = zsi - hi;
        T zl = zi - dx * (dzsi - dhi);
        T zn = zsn - hn;
        T zr = zn + dx * (dzsn - dhn);
        T zlr = dmax(zl, zr);
        T hl = hi - dx * dhi;
        T up = ui - dx * dui;
        T hp = dmax(T(0.0), hl + zl - zlr);
        T hr = hn + dx * dhn;
        T um = un + dx * dun;
        T hm = dmax(T(0.0), hr + zr - zlr);
        // KurgSolver
        T cp = A::sqrt_(C.g * hp);
        T cmo = A::sqrt_(C.g * hm);
        T ap = dmax(dmax(up + cp, um + cmo), T(0.0));
        T am = dmin(dmin(up - cp, um - cmo), T(0.0));
        T qm = hm * um;
        T qp = hp * up;
        T a = dmax(ap, -am);
        T fh, fu;
        if (a > C.epsi) {
            T ad = A::rcp(ap - am);
            T apm = ap * am;
            fh = (ap * qm - am * qp + apm * (hp - hm)) * ad;
            fu = (ap * (qm * um + C.ga * (hm * hm)) - am * (qp * up + C.ga * (hp * hp)) + apm * (qp - qm)) * ad;
            dt = A::div(C.CFL * C.delta, a);
        } else {
            fh = T(0.0);
            fu = T(0.0);
            dt = C.inv_epsi;
        }
        T fv = (fh > T(0.0) ? vn + dx * dvn : vi - dx * dvi) * fh;
        T sl = C.ga * (hp * hp - hl * hl + (hl + hi) * (zi - zl));
        T sr = C.ga * (hm * hm - hr * hr + (hr + hn) * (zn - zr));
        Fh = fh;
        Fq = fu - sl;
        S = fu - sr;
        Fv = fv;
        return true;
    }
    Fh = T(0.0);
    Fq = T(0.0);
    S = T(0.0);
    Fv = T(0.0);
    dt = C.inv_epsi;
    return false;
}

// Source cell of tile position (x, y) for a one-level grid: the block's own
// cell, the neighbour's cell, or (no neighbour) the block's own edge cell.
__device__ __forceinline__ int src_cell(int x, int y, int ib, int nL, int nR, int nB, int nT) {
    if (x < 0) return nL == ib ? memloc(0, y, ib) : memloc(BW + x, y, nL);
    if (x >= BW) return nR == ib ? memloc(BW - 1, y, ib) : memloc(x - BW, y, nR);
    if (y < 0) return nB == ib ? memloc(x, 0, ib) : memloc(x, BW + y, nB);
    if (y >= BW) return nT == ib ? memloc(x, BW - 1, ib) : memloc(x, y - BW, nT);
    return memloc(x, y, ib);
}

// The 128 strip positions (two deep on each side) of the 20x20 tile.
__device__ __forceinline__ void strip_pos(int k, int& x, int& y) {
    int s = k >> 4, j = k & 15;
    switch (s) {
        case 0: x = -1; y = j; break;
        case 1: x = -2; y = j; break;
        case 2: x = BW; y = j; break;
        case 3: x = BW + 1; y = j; break;
        case 4: x = j; y = -1; break;
        case 5: x = j; y = -2; break;
        case 6: x = j; y = BW; break;
        default: x = j; y = BW + 1; break;
    }
}

template <class T> struct Tile {
    T h[TP * TP], zs[TP * TP], u[TP * TP], v[TP * TP];
    T gx[4][16 * 18];  // d/dx of h, zs, u, v at x in [-1,16], y in [0,15]
    T gy[4][18 * 16];  // d/dy at x in [0,15], y in [-1,16]
    T fx[4][17 * 16];  // Fhu, Fqux, Su, Fqvx
    T fy[4][16 * 17];  // Fhv, Fqvy, Sv, Fquy
    T red[4];
};

// Gradients of the tile (interior pass by every thread, halo ring by wave 0).
template <class T, bool FAST> __device__ __forceinline__ void tile_gradients(Tile<T>& s, const Cst<T>& C, int t) {
    typedef Ar<T, FAST> A;
    const T th = C.theta;
    const T dl = C.delta;
    {
        int x = t & 15, y = t >> 4;
        int c = tix(x, y);
        const T* arr[4] = {s.h, s.zs, s.u, s.v};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const T* a = arr[k];
            s.gx[k][gxi(x, y)] = A::div(minmod2(th, a[c - 1], a[c], a[c + 1]), dl);
            s.gy[k][gyi(x, y)] = A::div(minmod2(th, a[c - TP], a[c], a[c + TP]), dl);
        }
    }
    if (t < 64) {
        int j = t & 15, side = t >> 4;
        const T* arr[4] = {s.h, s.zs, s.u, s.v};
        if (side < 2) {
            int x = side == 0 ? -1 : BW;
            int c = tix(x, j);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const T* a = arr[k];
                s.gx[k][gxi(x, j)] = A::div(minmod2(th, a[c - 1], a[c], a[c + 1]), dl);
            }
        } else {
            int y = side == 2 ? -1 : BW;
            int c = tix(j, y);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const T* a = arr[k];
                s.gy[k][gyi(j, y)] = A::div(minmod2(th, a[c - TP], a[c], a[c + TP]), dl);
            }
        }
    }
}

template <class T, bool FAST>
__device__ __forceinline__ bool xface(Tile<T>& s, const Cst<T>& C, int xf, int y, T& dt) {
    int ci = tix(xf, y), cn = ci - 1;
    int gi = gxi(xf, y), gn = gi - 1;
    T Fh, Fq, S, Fv;
    bool wet = kurg_face<T, FAST>(C, s.h[ci], s.h[cn], s.zs[ci], s.zs[cn], s.gx[0][gi], s.gx[0][gn], s.gx[1][gi],
                                  s.gx[1][gn], s.u[ci], s.u[cn], s.gx[2][gi], s.gx[2][gn], s.v[ci], s.v[cn],
                                  s.gx[3][gi], s.gx[3][gn], Fh, Fq, S, Fv, dt);
    int f = fxi(xf, y);
    s.fx[0][f] = Fh;
    s.fx[1][f] = Fq;
    s.fx[2][f] = S;
    s.fx[3][f] = Fv;
    return wet;
}

template <class T, bool FAST>
__device__ __forceinline__ bool yface(Tile<T>& s, const Cst<T>& C, int x, int yf, T& dt) {
    int ci = tix(x, yf), cn = ci - TP;
    int gi = gyi(x, yf), gn = gi - 16;
    T Fh, Fq, S, Fv;
    bool wet = kurg_face<T, FAST>(C, s.h[ci], s.h[cn], s.zs[ci], s.zs[cn], s.gy[0][gi], s.gy[0][gn], s.gy[1][gi],
                                  s.gy[1][gn], s.v[ci], s.v[cn], s.gy[3][gi], s.gy[3][gn], s.u[ci], s.u[cn],
                                  s.gy[2][gi], s.gy[2][gn], Fh, Fq, S, Fv, dt);
    int f = fyi(x, yf);
    s.fy[0][f] = Fh;
    s.fy[1][f] = Fq;
    s.fy[2][f] = S;
    s.fy[3][f] = Fv;
    return wet;
}

// Fluxes of the tile state, the wall patch and the tendencies of the
// thread's cell.  Returns the cell's dtmax (src/Kurganov.cu:102-105, 159).
template <class T, bool FAST>
__device__ __forceinline__ T tile_fluxes(Tile<T>& s, const Cst<T>& C, int t, bool wl, bool wr, bool wb, bool wt,
                                         T& dh, T& dhu, T& dhv) {
    tile_gradients<T, FAST>(s, C, t);
    __syncthreads();
    const int x = t & 15, y = t >> 4;
    T dtx, dty, dtd;
    bool wx = xface<T, FAST>(s, C, x, y, dtx);
    bool wy = yface<T, FAST>(s, C, x, y, dty);
    if (t < 16) xface<T, FAST>(s, C, BW, t, dtd);
    else if (t < 32) yface<T, FAST>(s, C, t - 16, BW, dtd);
    // X face then Y face write dtmax[i]; a dry face stores 1/epsi outright.
    T dtc = wx ? dtx : C.inv_epsi;
    if (wy) dtc = dty < dtc ? dty : dtc;
    else dtc = C.inv_epsi;
    __syncthreads();
    // Wall boundaries (noslipbndQ): left/bottom F=0 and S(inside)=G(face);
    // right/top F=0 and Fq(inside)=Ss(face).
    if (t < 16) {
        if (wl) { s.fx[0][fxi(0, t)] = T(0.0); s.fx[2][fxi(1, t)] = s.fx[1][fxi(0, t)]; }
        if (wr) { s.fx[0][fxi(BW, t)] = T(0.0); s.fx[1][fxi(BW - 1, t)] = s.fx[2][fxi(BW, t)]; }
    } else if (t < 32) {
        int j = t - 16;
        if (wb) { s.fy[0][fyi(j, 0)] = T(0.0); s.fy[2][fyi(j, 1)] = s.fy[1][fyi(j, 0)]; }
        if (wt) { s.fy[0][fyi(j, BW)] = T(0.0); s.fy[1][fyi(j, BW - 1)] = s.fy[2][fyi(j, BW)]; }
    }
    __syncthreads();
    // src/Advection.cu:236-250 (Cartesian: the metric terms vanish)
    const int f = fxi(x, y), g = fyi(x, y);
    const T hi = s.h[tix(x, y)];
    dh = T(-1.0) * (s.fx[0][f + 1] - s.fx[0][f] + s.fy[0][g + 16] - s.fy[0][g]) * C.cmdinv;
    dhu = (s.fx[1][f] + s.fy[3][g] - s.fx[2][f + 1] - s.fy[3][g + 16]) * C.cmdinv + C.fc * hi * s.v[tix(x, y)];
    dhv = (s.fy[1][g] + s.fx[3][f] - s.fy[2][g + 16] - s.fx[3][f + 1]) * C.cmdinv - C.fc * hi * s.u[tix(x, y)];
    return dtc;
}

// src/Advection.cu:332-385.  The drying limiter evaluates
// min(dt, max(h, eps)/|dh|) in double precision in the reference.
template <class T, bool FAST>
__device__ __forceinline__ void advance(const Cst<T>& C, int fmc, T dt, T hold, T uold, T vold, T dhi, T dhui, T dhvi,
                                        T& ho, T& uo, T& vo) {
    typedef Ar<T, FAST> A;
    T edt;
    if (fmc || dhi >= T(0.0)) {
        edt = dt;
    } else if (FAST && sizeof(T) == 4) {
        edt = dmin(dt, A::div(dmax(hold, C.eps), fabs(dhi)));
    } else {
        double lim = (hold > C.eps_d ? (double)hold : C.eps_d) / (double)fabs(dhi);
        edt = T((double)dt < lim ? (double)dt : lim);
    }
    ho = hold + edt * dhi;
    if (ho > C.eps) {
        T r = A::rcp(ho);
        if (FAST) {
            uo = (hold * uold + edt * dhui) * r;
            vo = (hold * vold + edt * dhvi) * r;
        } else {
            uo = (hold * uold + edt * dhui) / ho;
            vo = (hold * vold + edt * dhvi) / ho;
        }
    } else {
        uo = T(0.0);
        vo = T(0.0);
    }
}

template <class T> __device__ __forceinline__ T block_min(T v, T* red, int t) {
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    T m = red[0];
#pragma unroll
    for (int k = 1; k < 4; k++) m = red[k] < m ? red[k] : m;
    return m;
}

// Blocks beyond the step's end time skip all work: the device loop state
// carries a done flag.
__device__ __forceinline__ bool step_done(const DevLoop* L) { return L->done != 0; }

template <class T, bool FAST>
__global__ __launch_bounds__(256) void k_predictor(KParams P, StepBuffers<T> B, const DevLoop* L) {
    if (step_done(L)) return;
    __shared__ Tile<T> s;
    const int t = threadIdx.x;
    const int ibl = blockIdx.x;
    const int ib = P.active[ibl];
    const int nL = P.LeftBot[ib], nR = P.RightBot[ib], nB = P.BotLeft[ib], nT = P.TopLeft[ib];
    const Cst<T> C = make_cst<T, FAST>(P);
    {
        int x = t & 15, y = t >> 4;
        int src = memloc(x, y, ib);
        int c = tix(x, y);
        T hv = B.h[src];
        s.h[c] = hv;
        s.zs[c] = B.zb[src] + hv;
        s.u[c] = B.u[src];
        s.v[c] = B.v[src];
    }
    if (t < 128) {
        int x, y;
        strip_pos(t, x, y);
        int src = src_cell(x, y, ib, nL, nR, nB, nT);
        int c = tix(x, y);
        T hv = B.h[src];
        s.h[c] = hv;
        s.zs[c] = B.zb[src] + hv;
        s.u[c] = B.u[src];
        s.v[c] = B.v[src];
    }
    __syncthreads();
    const bool wall = P.aoibnd == 0;
    T dh, dhu, dhv;
    T dtc = tile_fluxes<T, FAST>(s, C, t, wall && nL == ib, wall && nR == ib, wall && nB == ib, wall && nT == ib,
                                 dh, dhu, dhv);
    const int i = memloc(t & 15, t >> 4, ib);
    B.dh[i] = dh;
    B.dhu[i] = dhu;
    B.dhv[i] = dhv;
    T m = block_min(dtc, s.red, t);
    if (t == 0) B.blockmin[ibl] = m;
}

template <class T, bool FAST>
__global__ __launch_bounds__(256) void k_corrector(KParams P, StepBuffers<T> B, DevLoop* L) {
    if (step_done(L)) return;
    __shared__ Tile<T> s;
    const int t = threadIdx.x;
    const int ibl = blockIdx.x;
    const int ib = P.active[ibl];
    const int nL = P.LeftBot[ib], nR = P.RightBot[ib], nB = P.BotLeft[ib], nT = P.TopLeft[ib];
    const Cst<T> C = make_cst<T, FAST>(P);
    const T dt = T(L->dt);
    const T hdt = dt * T(0.5);
    const int fmc = P.ForceMassConserve;
    // own cell: keep U^n for the full-step advance
    const int x0 = t & 15, y0 = t >> 4;
    const int i = memloc(x0, y0, ib);
    T h0 = B.h[i], u0 = B.u[i], v0 = B.v[i], zb0 = B.zb[i];
    {
        T ho, uo, vo;
        advance<T, FAST>(C, fmc, hdt, h0, u0, v0, B.dh[i], B.dhu[i], B.dhv[i], ho, uo, vo);
        int c = tix(x0, y0);
        s.h[c] = ho;
        s.zs[c] = zb0 + ho;
        s.u[c] = uo;
        s.v[c] = vo;
    }
    if (t < 128) {
        int x, y;
        strip_pos(t, x, y);
        int src = src_cell(x, y, ib, nL, nR, nB, nT);
        T ho, uo, vo;
        advance<T, FAST>(C, fmc, hdt, B.h[src], B.u[src], B.v[src], B.dh[src], B.dhu[src], B.dhv[src], ho, uo, vo);
        int c = tix(x, y);
        s.h[c] = ho;
        s.zs[c] = B.zb[src] + ho;
        s.u[c] = uo;
        s.v[c] = vo;
    }
    __syncthreads();
    const bool wall = P.aoibnd == 0;
    T dh, dhu, dhv;
    (void)tile_fluxes<T, FAST>(s, C, t, wall && nL == ib, wall && nR == ib, wall && nB == ib, wall && nT == ib,
                               dh, dhu, dhv);
    T h1, u1, v1;
    advance<T, FAST>(C, fmc, dt, h0, u0, v0, dh, dhu, dhv, h1, u1, v1);
    // bottom friction, src/Friction.cu:89-148
    if (h1 > C.eps) {
        typedef Ar<T, FAST> A;
        T normu = A::sqrt_(u1 * u1 + v1 * v1);
        T cfi;
        T cfv = B.cf[i];
        if (P.frictionmodel == 0) {
            cfi = cfv;
        } else if (P.frictionmodel == 1) {
            T Hbar = h1 / cfv;
            if (Hbar <= T(2.71828182845905)) cfi = T(1.0) / (T(0.46) * Hbar);
            else cfi = T(1.0) / (T(2.5) * (log(Hbar) - T(1.0) + T(1.359) / Hbar));
            cfi = cfi * cfi;
        } else {
            cfi = C.g * cfv * cfv / cbrt(h1);
        }
        T tb = cfi * normu / h1 * dt;
        u1 = u1 / (T(1.0) + tb);
        v1 = v1 / (T(1.0) + tb);
    }
    // implicit uniform rain after cleanup, src/Updateforcing.cu:530-581
    if (P.rain_on) {
        T act = zb0 < T(P.mask) ? T(1.0) : T(0.0);
        T rainhh = dmax(C.rainrate / T(1000.0) / T(3600.0) * dt, T(0.0)) * act;
        T qvol = h1 / (h1 + rainhh);
        bool wetold = (double)h1 > C.eps_d;
        h1 = h1 + rainhh;
        if (wetold) {
            u1 = u1 * qvol;
            v1 = v1 * qvol;
        }
    }
    B.ho[i] = h1;
    B.uo[i] = u1;
    B.vo[i] = v1;
}

// One workgroup: min of the block minima and 1/epsi, then the clamp and the
// output-time landing of CalctimestepCPU in FP64.
template <class T>
__global__ __launch_bounds__(1024) void k_reduce_dt(const T* __restrict__ blockmin, int nblk, DevLoop* L,
                                                    double endtime, int advance_time) {
    if (L->done) return;
    __shared__ T red[16];
    const int t = threadIdx.x;
    const T epsi = sizeof(T) == 8 ? T(2.220446049250313e-16) : T(1.1920928955078125e-07);
    T m = T(1.0) / epsi;
    for (int k = t; k < nblk; k += 1024) {
        T b = blockmin[k];
        m = b < m ? b : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(m, off, 64);
        m = o < m ? o : m;
    }
    if ((t & 63) == 0) red[t >> 6] = m;
    __syncthreads();
    if (t == 0) {
        T dt = red[0];
        for (int k = 1; k < 16; k++) dt = red[k] < dt ? red[k] : dt;
        if ((double)dt > 1.5 * L->dtmax) dt = T(1.5 * L->dtmax);
        double rem = L->nextoutputtime - L->totaltime;
        if (ceil(rem / (double)dt) > 0.0) dt = T(rem / ceil(rem / (double)dt));
        L->dt = (double)dt;
        L->dtmax = (double)dt;
        if (!isfinite((double)dt) || dt <= T(0)) L->bad = 1;
        if (advance_time) {
            L->totaltime = L->totaltime + L->dt;
            L->steps_done = L->steps_done + 1;
            if (L->totaltime >= endtime) L->done = 1;
        }
    }
}

// Per-cell initial step delta/sqrt(g*max(h,eps)) and its block minimum,
// src/Mainloop.cu:569-589.
template <class T>
__global__ __launch_bounds__(256) void k_init_dt(KParams P, const T* __restrict__ h, T* __restrict__ blockmin) {
    __shared__ T red[4];
    const int t = threadIdx.x;
    const int ib = P.active[blockIdx.x];
    const int i = memloc(t & 15, t >> 4, ib);
    T d = T(P.delta) / sqrt(T(P.g) * (h[i] > T(P.eps) ? h[i] : T(P.eps)));
    T m = block_min(d, red, t);
    if (t == 0) blockmin[blockIdx.x] = m;
}

// Block partial sums of h*delta^2 in FP64 (deterministic two-level sum).
template <class T>
__global__ __launch_bounds__(256) void k_volume(KParams P, const T* __restrict__ h, double* __restrict__ part) {
    __shared__ double red[4];
    const int t = threadIdx.x;
    const int ib = P.active[blockIdx.x];
    double v = (double)h[memloc(t & 15, t >> 4, ib)] * P.delta * P.delta;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    if (t == 0) part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ __launch_bounds__(1024) void k_sum(const double* __restrict__ part, int n, double* out) {
    __shared__ double red[16];
    const int t = threadIdx.x;
    double v = 0.0;
    for (int k = t; k < n; k += 1024) v += part[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((t & 63) == 0) red[t >> 6] = v;
    __syncthreads();
    if (t == 0) {
        double s2 = 0.0;
        for (int k = 0; k < 16; k++) s2 += red[k];
        *out = s2;
    }
}

// cleanupGPU semantics (src/Advection.cu:399-417): copy interior h, u, v.
template <class T>
__global__ __launch_bounds__(256) void k_copy_state(KParams P, const T* __restrict__ h, const T* __restrict__ u,
                                                    const T* __restrict__ v, T* __restrict__ ho, T* __restrict__ uo,
                                                    T* __restrict__ vo) {
    const int t = threadIdx.x;
    const int ib = P.active[blockIdx.x];
    const int i = memloc(t & 15, t >> 4, ib);
    ho[i] = h[i];
    uo[i] = u[i];
    vo[i] = v[i];
}

// ------------------------------------------------------------------ launch
template <class T, bool FAST>
void launch_step_t(const KParams& P, const StepBuffers<T>& B, DevLoop* L, double endtime, int advance_time,
                   hipStream_t st, hipEvent_t* ev) {
    if (ev) hipEventRecord(ev[0], st);
    hipLaunchKernelGGL((k_predictor<T, FAST>), dim3(P.nblk), dim3(256), 0, st, P, B, L);
    if (ev) hipEventRecord(ev[1], st);
    hipLaunchKernelGGL((k_reduce_dt<T>), dim3(1), dim3(1024), 0, st, B.blockmin, P.nblk, L, endtime, advance_time);
    if (ev) hipEventRecord(ev[2], st);
    hipLaunchKernelGGL((k_corrector<T, FAST>), dim3(P.nblk), dim3(256), 0, st, P, B, L);
    if (ev) hipEventRecord(ev[3], st);
}

void launch_step(int precision, const KParams& P, const void* buffers, DevLoop* L, double endtime, int advance_time,
                 hipStream_t st, hipEvent_t* ev) {
    if (precision == 0)
        launch_step_t<double, false>(P, *(const StepBuffers<double>*)buffers, L, endtime, advance_time, st, ev);
    else if (precision == 1)
        launch_step_t<float, false>(P, *(const StepBuffers<float>*)buffers, L, endtime, advance_time, st, ev);
    else
        launch_step_t<float, true>(P, *(const StepBuffers<float>*)buffers, L, endtime, advance_time, st, ev);
}

void launch_init_dt(int precision, const KParams& P, const void* h, void* blockmin, DevLoop* L, hipStream_t st) {
    if (precision == 0) {
        hipLaunchKernelGGL(k_init_dt<double>, dim3(P.nblk), dim3(256), 0, st, P, (const double*)h, (double*)blockmin);
        hipLaunchKernelGGL(k_reduce_dt<double>, dim3(1), dim3(1024), 0, st, (const double*)blockmin, P.nblk, L, 0.0, 0);
    } else {
        hipLaunchKernelGGL(k_init_dt<float>, dim3(P.nblk), dim3(256), 0, st, P, (const float*)h, (float*)blockmin);
        hipLaunchKernelGGL(k_reduce_dt<float>, dim3(1), dim3(1024), 0, st, (const float*)blockmin, P.nblk, L, 0.0, 0);
    }
}

void launch_volume(int precision, const KParams& P, const void* h, double* part, double* out, hipStream_t st) {
    if (precision == 0) hipLaunchKernelGGL(k_volume<double>, dim3(P.nblk), dim3(256), 0, st, P, (const double*)h, part);
    else hipLaunchKernelGGL(k_volume<float>, dim3(P.nblk), dim3(256), 0, st, P, (const float*)h, part);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(1024), 0, st, part, P.nblk, out);
}

void launch_copy_state(int precision, const KParams& P, const void* const* src, void* const* dst, hipStream_t st) {
    if (precision == 0)
        hipLaunchKernelGGL(k_copy_state<double>, dim3(P.nblk), dim3(256), 0, st, P, (const double*)src[0],
                           (const double*)src[1], (const double*)src[2], (double*)dst[0], (double*)dst[1],
                           (double*)dst[2]);
    else
        hipLaunchKernelGGL(k_copy_state<float>, dim3(P.nblk), dim3(256), 0, st, P, (const float*)src[0],
                           (const float*)src[1], (const float*)src[2], (float*)dst[0], (float*)dst[1],
                           (float*)dst[2]);
}

}  // namespace bgf
