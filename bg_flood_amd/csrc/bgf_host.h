// bgf_host.h -- host-side helpers shared by the C ABI translation units:
// error reporting, parameter checks and the processing order of the
// persistent kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/bgflood_amd.h"
#include "bgf_kernels.h"

namespace bgf_host {

constexpr int BW = 16;
constexpr int MW = 18;
constexpr int BS = MW * MW;

inline thread_local std::string g_err;

inline bgf_status fail(bgf_status s, const std::string& msg) {
    g_err = msg;
    return s;
}

#define HIPCHK(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return fail(BGF_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline int memloc(int i, int j, int ib) { return (i + 1) + (j + 1) * MW + ib * BS; }

inline double level_dx(double dx, int lev) { return lev < 0 ? dx * (1 << -lev) : dx / (1 << lev); }

inline bgf_status check_param(const bgf_param* p) {
    if (!p) return fail(BGF_ERR_INVALID_ARG, "null param");
    if (p->engine != 2) return fail(BGF_ERR_UNSUPPORTED, "only engine 2 (Kurganov) is provided");
    if (p->spherical) return fail(BGF_ERR_UNSUPPORTED, "spherical grids are not provided");
    if (p->nblk <= 0 || p->nblkmem < p->nblk) return fail(BGF_ERR_INVALID_ARG, "need 0 < nblk <= nblkmem");
    if (p->precision < 0 || p->precision > 3) return fail(BGF_ERR_INVALID_ARG, "bad precision");
    if (p->aoibnd != 0 && p->aoibnd != 1)
        return fail(BGF_ERR_UNSUPPORTED, "remainder boundary type must be 0 (wall) or 1 (Neumann)");
    if (p->frictionmodel < -1 || p->frictionmodel > 1) return fail(BGF_ERR_INVALID_ARG, "bad frictionmodel");
    if (!(p->delta > 0.0) || !(p->g > 0.0) || !(p->eps > 0.0)) return fail(BGF_ERR_INVALID_ARG, "bad g/eps/delta");
    return BGF_OK;
}

template <class T> inline double coriolis(double lat) {
    // src/Advection.cu:187 (Cartesian): sin(T(lat*pi/180))*pi/T(21600) with pi = 3.14159265
    const double pi = 3.14159265;
    return (double)T(std::sin(T(lat * pi / 180.0)) * pi / T(21600.0));
}

// Processing order of the persistent kernels: the active blocks grouped in
// tiles of TILE_X x TILE_Y blocks (row-major tiles, row-major blocks inside a
// tile), by their integer block coordinates.  Each XCD processes consecutive
// positions of this order, so the blocks it has in flight at once form a
// compact patch whose neighbours are mostly in its own L2.
constexpr int TILE_X = 16, TILE_Y = 12;
inline std::vector<int> processing_order(const bgf_blocks* b, int nblk, double dx) {
    std::vector<int> order(b->active, b->active + nblk);
    std::vector<long long> key(nblk);
    long long maxbx = 1;
    std::vector<long long> bx(nblk), by(nblk);
    for (int k = 0; k < nblk; k++) {
        const int ib = b->active[k];
        const int lev = b->level[ib];
        const double levdx = lev < 0 ? dx * (1 << -lev) : dx / (1 << lev);
        bx[k] = (long long)std::floor(b->xo[ib] / (BW * levdx));
        by[k] = (long long)std::floor(b->yo[ib] / (BW * levdx));
        if (bx[k] + 1 > maxbx) maxbx = bx[k] + 1;
    }
    const long long ntx = (maxbx + TILE_X - 1) / TILE_X;
    for (int k = 0; k < nblk; k++) {
        const long long tx = bx[k] / TILE_X, ty = by[k] / TILE_Y;
        key[k] = ((ty * ntx + tx) * TILE_Y + by[k] % TILE_Y) * TILE_X + bx[k] % TILE_X;
    }
    std::vector<int> idx(nblk);
    for (int k = 0; k < nblk; k++) idx[k] = k;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int c) { return key[a] < key[c]; });
    for (int k = 0; k < nblk; k++) order[k] = b->active[idx[k]];
    return order;
}

}  // namespace bgf_host
