// bgf_adapt.cpp -- one adaptation pass of the block quadtree, with the block
// numbering of the reference's adaptation, so that a mesh adapted here has the
// same block indices, active order, levels and origins as one adapted by
// BG_Flood (src/Adaptation.cu:17-1879).
//
//   refinesanitycheck  src/Adaptation.cu:157-296: no refinement at maxlevel,
//                      no coarsening at minlevel; a refining block keeps its
//                      eight neighbours from coarsening and forces its
//                      coarser direct neighbours (TopLeft, BotLeft, LeftBot,
//                      RightBot) to refine, to a fixed point
//                      (checkneighbourrefine, :298-325); a block coarsens only
//                      as the bottom-left block of four same-level siblings
//                      that all coarsen, none beside a finer block.  The flags
//                      are updated in place in active-list order, as there.
//                      The reference then cancels every coarsening (:170-175);
//                      bgf_adapt_refine keeps that, bgf_adapt_mesh applies the
//                      coarsening when coarsen flags are given.
//   CalcAvailblk       :535-598: the k-th refining block in active order takes
//                      free block slots 3k..3k+2 (free = not active, ascending,
//                      before this pass's coarsening frees any).
//   AddBlocks          :608-662: if fewer slots are free than needed,
//                      nblkmem = ceil((nblk + 3*nrefine) * 1.05).
//   coarsen            :677-1004: the bottom-left block becomes the parent
//                      (origin moved by half a fine cell), the other three
//                      leave the active list.
//   refine             :1017-1795: the block keeps its index as the bottom-left
//                      child; bottom-right, top-left and top-right take the
//                      free slots; origins move by half a child cell; the new
//                      blocks are appended to the active list in refining order.
//   Adaptationcleanup  :1807-1876: levels of active blocks updated, the active
//                      list compacted in order.
// Neighbours are then rebuilt from the block positions (the reference updates
// them incrementally; the result is the geometric neighbour relation).
//
// The remap table says where each slot's state comes from, for bgf_remesh:
// kind -1 inactive, 0 kept, 1 child (src0 = parent, src1 = quadrant 0 BL,
// 1 BR, 2 TL, 3 TR), 2 coarsened (src0..src3 = the old BL, BR, TL, TR blocks).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "bgf_host.h"

using namespace bgf_host;

namespace {

template <class T> double stored(double v) { return (double)(T)v; }

struct MeshOut {
    int *active, *level, *nb[8];
    double *xo, *yo;
};

bgf_status adapt_mesh(const bgf_param* p, const bgf_blocks* b, const unsigned char* refine_flags,
                      const unsigned char* coarsen_flags, int nbx, int nby, int capacity, int* nblk_out,
                      int* nblkmem_out, const MeshOut& o, int* remap) {
    if (!p || !b || !refine_flags || !nblk_out || !nblkmem_out || !o.active || !o.level || !o.xo || !o.yo)
        return fail(BGF_ERR_INVALID_ARG, "null argument");
    for (int k = 0; k < 8; k++)
        if (!o.nb[k]) return fail(BGF_ERR_INVALID_ARG, "null argument");
    const int nblk = p->nblk, nbm = p->nblkmem;
    if (nblk <= 0 || nbm < nblk || nbx <= 0 || nby <= 0) return fail(BGF_ERR_INVALID_ARG, "bad sizes");
    const bool dbl = p->precision == BGF_FP64;
    auto st = [dbl](double v) { return dbl ? stored<double>(v) : stored<float>(v); };
    const int* LB = b->LeftBot;
    const int* LT = b->LeftTop;
    const int* RB = b->RightBot;
    const int* RT = b->RightTop;
    const int* TL = b->TopLeft;
    const int* TR = b->TopRight;
    const int* BL = b->BotLeft;
    const int* BR = b->BotRight;
    std::vector<int> lev(b->level, b->level + nbm);
    std::vector<char> refine(nbm, 0), coarsen(nbm, 0), act(nbm, 0);
    for (int k = 0; k < nblk; k++) {
        const int ib = b->active[k];
        if (ib < 0 || ib >= nbm || act[ib]) return fail(BGF_ERR_INVALID_ARG, "bad active list");
        act[ib] = 1;
        refine[ib] = refine_flags[ib] != 0;
        coarsen[ib] = coarsen_flags ? coarsen_flags[ib] != 0 : 0;
    }
    for (int k = 0; k < nblk; k++) {
        const int ib = b->active[k];
        for (const int* a : {LB, LT, RB, RT, TL, TR, BL, BR})
            if (a[ib] < 0 || a[ib] >= nbm || !act[a[ib]]) return fail(BGF_ERR_INVALID_ARG, "bad neighbour");
    }
    // refinesanitycheck
    for (int k = 0; k < nblk; k++) {
        const int ib = b->active[k];
        if (refine[ib] && lev[ib] == p->maxlevel) refine[ib] = 0;
        if (coarsen[ib] && lev[ib] == p->minlevel) coarsen[ib] = 0;
    }
    for (int k = 0; k < nblk; k++) {
        const int ib = b->active[k];
        if (refine[ib])
            for (const int* a : {RB, RT, LB, LT, TL, TR, BL, BR}) coarsen[a[ib]] = 0;
    }
    for (int k = 0; k < nblk; k++) {
        const int ib = b->active[k];
        const int l = lev[ib];
        if (coarsen[ib] && (l < lev[LB[ib]] || l < lev[RB[ib]] || l < lev[TL[ib]] || l < lev[BL[ib]])) coarsen[ib] = 0;
    }
    for (int iter = 1; iter > 0;) {
        iter = 0;
        for (int k = 0; k < nblk; k++) {
            const int ib = b->active[k];
            if (!refine[ib]) continue;
            for (int n : {TL[ib], BL[ib], LB[ib], RB[ib]}) {
                if (!refine[n] && lev[n] < lev[ib]) {
                    refine[n] = 1;
                    coarsen[n] = 0;
                    iter++;
                }
                if (lev[n] == lev[ib]) coarsen[n] = 0;
            }
        }
    }
    for (int k = 0; k < nblk; k++) {
        const int ib = b->active[k];
        if (!coarsen[ib]) continue;
        const double d = level_dx(p->dx, lev[ib]);
        const int r = RB[ib], t = TL[ib], tr = RB[TR[ib]];
        const bool corner = (int)(b->xo[ib] / d / BW) % 2 == 0 && (int)(b->yo[ib] / d / BW) % 2 == 0 && r != ib &&
                            t != ib && tr != TR[ib];
        const bool same = corner && lev[ib] == lev[r] && lev[ib] == lev[t] && lev[ib] == lev[tr];
        if (!(same && coarsen[r] && coarsen[t] && coarsen[tr])) coarsen[ib] = 0;
    }
    // CalcAvailblk / AddBlocks
    int nrefine = 0;
    std::vector<int> csum(nbm, 0);
    for (int k = 0, c = -3; k < nblk; k++) {
        const int ib = b->active[k];
        if (refine[ib]) {
            nrefine++;
            c += 3;
        }
        csum[ib] = c;
    }
    const int nnew = 3 * nrefine;
    int nbm2 = nbm;
    int navail = 0;
    for (int ib = 0; ib < nbm; ib++) navail += act[ib] ? 0 : 1;
    if (nnew > navail) nbm2 = (int)std::ceil((nblk + nnew) * 1.05);
    if (nbm2 > capacity) return fail(BGF_ERR_OOM, "adapted mesh exceeds the output capacity");
    std::vector<int> avail;
    for (int ib = 0; ib < nbm2; ib++)
        if (ib >= nbm || !act[ib]) avail.push_back(ib);
    std::vector<double> x(nbm2, 0.0), y(nbm2, 0.0);
    std::vector<int> nl(nbm2, 0), kind(nbm2, -1), src(4 * (size_t)nbm2, -1);
    for (int ib = 0; ib < nbm; ib++) {
        x[ib] = b->xo[ib];
        y[ib] = b->yo[ib];
        nl[ib] = lev[ib];
    }
    std::vector<char> gone(nbm2, 0);
    // coarsen (src/Adaptation.cu:677-1004; origins stored in T)
    for (int k = 0; k < nblk; k++) {
        const int ib = b->active[k];
        if (!coarsen[ib]) continue;
        const int r = RB[ib], t = TL[ib], tr = TL[RB[ib]];
        const double half = level_dx(p->dx, lev[ib] + 1);
        x[ib] = st(x[ib] + st(half));
        y[ib] = st(y[ib] + st(half));
        nl[ib] = lev[ib] - 1;
        gone[r] = gone[t] = gone[tr] = 1;
        kind[ib] = 2;
        src[4 * ib] = ib;
        src[4 * ib + 1] = r;
        src[4 * ib + 2] = t;
        src[4 * ib + 3] = tr;
    }
    // refine (src/Adaptation.cu:1118-1160)
    for (int k = 0; k < nblk; k++) {
        const int ib = b->active[k];
        if (!refine[ib]) continue;
        const int kb[4] = {ib, avail[csum[ib]], avail[csum[ib] + 1], avail[csum[ib] + 2]};
        const double delx = level_dx(p->dx, lev[ib] + 1);
        const double xob = x[ib] - 0.5 * delx, yob = y[ib] - 0.5 * delx;
        x[kb[0]] = st(xob);
        y[kb[0]] = st(yob);
        x[kb[1]] = st(xob + BW * delx);
        y[kb[1]] = st(yob);
        x[kb[2]] = st(xob);
        y[kb[2]] = st(yob + BW * delx);
        x[kb[3]] = st(xob + BW * delx);
        y[kb[3]] = st(yob + BW * delx);
        for (int q = 0; q < 4; q++) {
            nl[kb[q]] = lev[ib] + 1;
            kind[kb[q]] = 1;
            src[4 * kb[q]] = ib;
            src[4 * kb[q] + 1] = q;
        }
    }
    // Adaptationcleanup: the kept active order, then the new blocks in refining order
    std::vector<int> order;
    for (int k = 0; k < nblk; k++)
        if (!gone[b->active[k]]) order.push_back(b->active[k]);
    for (int k = 0; k < nblk; k++) {
        const int ib = b->active[k];
        if (!refine[ib]) continue;
        for (int j = 0; j < 3; j++) order.push_back(avail[csum[ib] + j]);
    }
    for (int ib : order)
        if (kind[ib] < 0) {
            kind[ib] = 0;
            src[4 * ib] = ib;
        }
    const int n2 = (int)order.size();
    // neighbours from the positions
    std::vector<int> L(n2);
    std::vector<long long> ix(n2), iy(n2);
    for (int k = 0; k < n2; k++) {
        const int ib = order[k];
        const double ldx = level_dx(p->dx, nl[ib]);
        L[k] = nl[ib];
        ix[k] = std::llround((x[ib] - 0.5 * ldx) / (BW * ldx));
        iy[k] = std::llround((y[ib] - 0.5 * ldx) / (BW * ldx));
    }
    for (int ib = 0; ib < nbm2; ib++) {  // unused slots point at themselves
        for (int s = 0; s < 8; s++) o.nb[s][ib] = ib;
        o.active[ib] = -1;
        o.level[ib] = nl[ib];
        o.xo[ib] = x[ib];
        o.yo[ib] = y[ib];
    }
    // the domain is nbx x nby blocks of level 0 (InitMesh at initlevel 0)
    bgf_status s = quadtree_neighbours(n2, order.data(), L.data(), ix.data(), iy.data(), nbx, nby, 0, o.nb[0], o.nb[1],
                                       o.nb[2], o.nb[3], o.nb[4], o.nb[5], o.nb[6], o.nb[7]);
    if (s != BGF_OK) return s;
    for (int k = 0; k < n2; k++) o.active[k] = order[k];
    if (remap)
        for (int ib = 0; ib < nbm2; ib++) {
            remap[5 * ib] = kind[ib];
            for (int q = 0; q < 4; q++) remap[5 * ib + 1 + q] = src[4 * ib + q];
        }
    *nblk_out = n2;
    *nblkmem_out = nbm2;
    return BGF_OK;
}

}  // namespace

extern "C" {

bgf_status bgf_adapt_refine(const bgf_param* p, const bgf_blocks* b, const unsigned char* refine_flags, int nbx,
                            int nby, int capacity, int* nblk_out, int* nblkmem_out, int* active, int* level,
                            double* xo, double* yo, int* LeftBot, int* LeftTop, int* RightBot, int* RightTop,
                            int* TopLeft, int* TopRight, int* BotLeft, int* BotRight) {
    const MeshOut o = {active, level, {LeftBot, LeftTop, RightBot, RightTop, TopLeft, TopRight, BotLeft, BotRight},
                       xo, yo};
    return adapt_mesh(p, b, refine_flags, nullptr, nbx, nby, capacity, nblk_out, nblkmem_out, o, nullptr);
}

bgf_status bgf_adapt_mesh(const bgf_param* p, const bgf_blocks* b, const unsigned char* refine_flags,
                          const unsigned char* coarsen_flags, int nbx, int nby, int capacity, int* nblk_out,
                          int* nblkmem_out, const bgf_blocks_out* out, int* remap) {
    if (!out) return fail(BGF_ERR_INVALID_ARG, "null argument");
    const MeshOut o = {out->active,
                       out->level,
                       {out->LeftBot, out->LeftTop, out->RightBot, out->RightTop, out->TopLeft, out->TopRight,
                        out->BotLeft, out->BotRight},
                       out->xo,
                       out->yo};
    return adapt_mesh(p, b, refine_flags, coarsen_flags, nbx, nby, capacity, nblk_out, nblkmem_out, o, remap);
}

}  // extern "C"
