/*
 * bgflood_amd.hpp -- C++ drop-in for BG_Flood's FlowGPU<T> / FlowCPU<T>
 * (src/FlowGPU.cu:14-315, src/FlowCPU.cu:15-253) over the C ABI of
 * bgflood_amd.h.
 *
 * The templates read BG_Flood's own types by field name -- Param
 * (src/Param.h:10-239), Loop<T> (src/Arrays.h:358-366), Model<T> with its
 * BlockP<T> blocks, EvolvingP<T> evolv, zb and cf (src/Arrays.h:13-404) --
 * so BG_Flood's MainLoop can keep calling FlowGPU(XParam, XLoop, XForcing,
 * XModel) with this header in place of src/FlowGPU.h:
 *
 *     bgflood_amd::Flow<T> flow;                    // one per domain
 *     flow.init(XParam, XModel, rain_mmh);          // after SetupGPU's host arrays exist
 *     flow.step(XLoop);                             // FlowGPU<T>(XParam, XLoop, XForcing, XModel)
 *     flow.download(XModel);                        // before outputs
 *
 * or the function form FlowGPU(flow, XLoop).  Only the arrays travel; no
 * BG_Flood header is needed to compile this file.
 */
#ifndef BGFLOOD_AMD_HPP
#define BGFLOOD_AMD_HPP

#include <cstddef>
#include <vector>

#include "bgflood_amd.h"

namespace bgflood_amd {

/* bgf_param from a BG_Flood Param (the fields the step reads). */
template <class Param> bgf_param param_from(const Param& P, int precision, int nblk, int nblkmem, double rain_mmh) {
    bgf_param p = {};
    p.g = P.g;
    p.eps = P.eps;
    p.CFL = P.CFL;
    p.theta = P.theta;
    p.dx = P.dx;
    p.delta = P.delta;
    p.xo = P.xo;
    p.yo = P.yo;
    p.lat = P.lat;
    p.mask = P.mask;
    p.frictionmodel = P.frictionmodel;
    p.engine = P.engine;
    p.aoibnd = P.aoibnd;
    p.ForceMassConserve = P.ForceMassConserve ? 1 : 0;
    p.wetdryfix = P.wetdryfix ? 1 : 0;
    p.spherical = P.spherical ? 1 : 0;
    p.nblk = nblk;
    p.nblkmem = nblkmem;
    p.minlevel = P.minlevel;
    p.maxlevel = P.maxlevel;
    p.rain_on = rain_mmh != 0.0 ? 1 : 0;
    p.rain = rain_mmh;
    p.precision = precision;
    p.Radius = P.Radius;
    return p;
}

template <class T> class Flow {
  public:
    Flow() = default;
    Flow(const Flow&) = delete;
    Flow& operator=(const Flow&) = delete;
    ~Flow() {
        if (h_) bgf_destroy(h_);
    }

    /* Allocates the device Model<T> and uploads the state: AllocateGPU +
     * CopytoGPU of SetupGPU (src/Setup_GPU.cu:15-157).  precision defaults to
     * the reference arithmetic of T. */
    template <class Param, class Model>
    bgf_status init(const Param& XParam, const Model& XModel, double rain_mmh = 0.0, int precision = -1) {
        if (precision < 0) precision = sizeof(T) == sizeof(double) ? BGF_FP64 : BGF_FP32;
        const int nbm = XParam.nblkmem;
        bgf_param p = param_from(XParam, precision, XParam.nblk, nbm, rain_mmh);
        xo_.assign(nbm, 0.0);
        yo_.assign(nbm, 0.0);
        for (int i = 0; i < nbm; i++) {
            xo_[i] = (double)XModel.blocks.xo[i];
            yo_[i] = (double)XModel.blocks.yo[i];
        }
        const auto& B = XModel.blocks;
        bgf_blocks b = {B.active,   B.level,    B.LeftBot,  B.LeftTop, B.RightBot, B.RightTop, B.TopLeft,
                        B.TopRight, B.BotLeft, B.BotRight, xo_.data(), yo_.data()};
        if (h_) {
            bgf_destroy(h_);
            h_ = nullptr;
        }
        bgf_status s = bgf_create(&p, &b, &h_);
        if (s != BGF_OK) return s;
        return bgf_set_state(h_, XModel.zb, XModel.evolv.h, XModel.evolv.u, XModel.evolv.v, XModel.cf);
    }

    /* FlowGPU<T>(XParam, XLoop, XForcing, XModel): one step; writes
     * XLoop.dt and XLoop.dtmax; the caller advances XLoop.totaltime
     * (src/Mainloop.cu:60). */
    template <class Loop> bgf_status step(Loop& XLoop) {
        bgf_loop l = {(double)XLoop.nextoutputtime, (double)XLoop.dt, (double)XLoop.dtmax, (double)XLoop.totaltime};
        bgf_status s = bgf_step(h_, &l);
        if (s != BGF_OK) return s;
        XLoop.dt = l.dt;
        XLoop.dtmax = l.dtmax;
        return BGF_OK;
    }

    /* Runs to the loop's next output time on the GPU (steps and time keeping
     * stay on the device), as MainLoop's inner loop between outputs. */
    template <class Loop> bgf_status run_to_output(Loop& XLoop, int max_steps = 1 << 30) {
        bgf_loop l = {(double)XLoop.nextoutputtime, (double)XLoop.dt, (double)XLoop.dtmax, (double)XLoop.totaltime};
        bgf_status s = bgf_run(h_, &l, max_steps, l.nextoutputtime);
        if (s != BGF_OK) return s;
        XLoop.dt = l.dt;
        XLoop.dtmax = l.dtmax;
        XLoop.totaltime = l.totaltime;
        return BGF_OK;
    }

    /* Initial time step, initdt (src/Mainloop.cu:515-554). */
    template <class Loop> bgf_status init_dt(Loop& XLoop, double dtinit = -1.0) {
        bgf_loop l = {(double)XLoop.nextoutputtime, (double)XLoop.dt, (double)XLoop.dtmax, (double)XLoop.totaltime};
        bgf_status s = bgf_init_dt(h_, dtinit, &l);
        if (s != BGF_OK) return s;
        XLoop.dtmax = l.dtmax;
        return BGF_OK;
    }

    /* CopyGPUtoCPU of the evolving state before outputs. */
    template <class Model> bgf_status download(Model& XModel) {
        return bgf_get_state(h_, XModel.evolv.zs, XModel.evolv.h, XModel.evolv.u, XModel.evolv.v);
    }

    bgf_handle handle() const { return h_; }

  private:
    bgf_handle h_ = nullptr;
    std::vector<double> xo_, yo_;
};

/* Function form of the reference's entry point. */
template <class T, class Loop> bgf_status FlowGPU(Flow<T>& flow, Loop& XLoop) { return flow.step(XLoop); }

}  // namespace bgflood_amd

#endif
