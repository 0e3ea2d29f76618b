// bgf_adapt.cpp -- one refinement pass of the block quadtree, with the
// block numbering of the reference's adaptation, so that a mesh adapted here
// has the same block indices, active order, levels and origins as one adapted
// by BG_Flood.
//
// The reference cancels every coarsening request (src/Adaptation.cu:170-175),
// so adaptation only refines:
//   refinesanitycheck  src/Adaptation.cu:157-296: no refinement at maxlevel;
//                      a refining block forces its coarser direct neighbours
//                      (TopLeft, BotLeft, LeftBot, RightBot) to refine, to a
//                      fixed point (checkneighbourrefine, :298-325).
//   CalcAvailblk       :535-598: the k-th refining block in active order takes
//                      free block slots 3k..3k+2 (free = not active, ascending).
//   AddBlocks          :608-662: if fewer slots are free than needed,
//                      nblkmem = ceil((nblk + 3*nrefine) * 1.05).
//   refine             :1017-1795: the block keeps its index as the bottom-left
//                      child; bottom-right, top-left and top-right take the
//                      free slots; origins move by half a child cell; the new
//                      blocks are appended to the active list in refining order.
// Neighbours are then rebuilt from the block positions (the reference updates
// them incrementally; the result is the geometric neighbour relation).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "bgf_host.h"

using namespace bgf_host;

namespace {

template <class T> double stored(double v) { return (double)(T)v; }

}  // namespace


extern "C" {

bgf_status bgf_adapt_refine(const bgf_param* p, const bgf_blocks* b, const unsigned char* refine_flags, int nbx,
                            int nby, int capacity, int* nblk_out, int* nblkmem_out, int* active, int* level,
                            double* xo, double* yo, int* LeftBot, int* LeftTop, int* RightBot, int* RightTop,
                            int* TopLeft, int* TopRight, int* BotLeft, int* BotRight) {
    if (!p || !b || !refine_flags || !nblk_out || !nblkmem_out || !active || !level || !xo || !yo)
        return fail(BGF_ERR_INVALID_ARG, "null argument");
    const int nblk = p->nblk, nbm = p->nblkmem;
    if (nblk <= 0 || nbm < nblk || nbx <= 0 || nby <= 0) return fail(BGF_ERR_INVALID_ARG, "bad sizes");
    const bool dbl = p->precision == BGF_FP64;
    std::vector<int> lev(b->level, b->level + nbm);
    std::vector<char> refine(nbm, 0), act(nbm, 0);
    for (int k = 0; k < nblk; k++) {
        const int ib = b->active[k];
        if (ib < 0 || ib >= nbm || act[ib]) return fail(BGF_ERR_INVALID_ARG, "bad active list");
        act[ib] = 1;
        refine[ib] = refine_flags[ib] != 0;
    }
    // refinesanitycheck (coarsening is cancelled by the reference)
    for (int k = 0; k < nblk; k++) {
        const int ib = b->active[k];
        if (refine[ib] && lev[ib] == p->maxlevel) refine[ib] = 0;
    }
    for (int iter = 1; iter > 0;) {
        iter = 0;
        for (int k = 0; k < nblk; k++) {
            const int ib = b->active[k];
            if (!refine[ib]) continue;
            for (int n : {b->TopLeft[ib], b->BotLeft[ib], b->LeftBot[ib], b->RightBot[ib]})
                if (!refine[n] && lev[n] < lev[ib]) {
                    refine[n] = 1;
                    iter++;
                }
        }
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
    // refine: levels and origins (src/Adaptation.cu:1118-1160, origins stored in T)
    std::vector<double> x(nbm2, 0.0), y(nbm2, 0.0);
    std::vector<int> nl(nbm2, 0);
    for (int ib = 0; ib < nbm; ib++) {
        x[ib] = b->xo[ib];
        y[ib] = b->yo[ib];
        nl[ib] = lev[ib];
    }
    std::vector<int> order(b->active, b->active + nblk);
    for (int k = 0; k < nblk; k++) {
        const int ib = b->active[k];
        if (!refine[ib]) continue;
        const int br = avail[csum[ib]], tl = avail[csum[ib] + 1], tr = avail[csum[ib] + 2];
        const double delx = level_dx(p->dx, lev[ib] + 1);
        const double xob = x[ib] - 0.5 * delx, yob = y[ib] - 0.5 * delx;
        auto st = [dbl](double v) { return dbl ? stored<double>(v) : stored<float>(v); };
        x[ib] = st(xob);
        y[ib] = st(yob);
        x[br] = st(xob + BW * delx);
        y[br] = st(yob);
        x[tl] = st(xob);
        y[tl] = st(yob + BW * delx);
        x[tr] = st(xob + BW * delx);
        y[tr] = st(yob + BW * delx);
        for (int c : {ib, br, tl, tr}) nl[c] = lev[ib] + 1;
    }
    for (int k = 0; k < nblk; k++) {  // activation order (:1766-1789)
        const int ib = b->active[k];
        if (!refine[ib]) continue;
        for (int j = 0; j < 3; j++) order.push_back(avail[csum[ib] + j]);
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
        LeftBot[ib] = LeftTop[ib] = RightBot[ib] = RightTop[ib] = ib;
        TopLeft[ib] = TopRight[ib] = BotLeft[ib] = BotRight[ib] = ib;
        active[ib] = -1;
        level[ib] = nl[ib];
        xo[ib] = x[ib];
        yo[ib] = y[ib];
    }
    // the domain is nbx x nby blocks of level 0 (InitMesh at initlevel 0)
    bgf_status s = quadtree_neighbours(n2, order.data(), L.data(), ix.data(), iy.data(), nbx, nby, 0, LeftBot, LeftTop,
                                       RightBot, RightTop, TopLeft, TopRight, BotLeft, BotRight);
    if (s != BGF_OK) return s;
    for (int k = 0; k < n2; k++) active[k] = order[k];
    *nblk_out = n2;
    *nblkmem_out = nbm2;
    return BGF_OK;
}

}  // extern "C"
