// Builds the by-value kernel arguments from a context.
#pragma once

#include <cstring>

#include "context.h"
#include "fiber_ops.h"
#include "solid_ops.h"

namespace cymg {

// The argument structs are zeroed first so that their padding bytes are
// defined: the captured iteration graphs are keyed on the argument bytes.
inline FiberArgs make_fiber_args(const Context& c)
{
    FiberArgs a;
    std::memset(&a, 0, sizeof(a));
    a.np = c.d_np.get();
    a.idx = c.d_idx.get();
    a.seg = c.d_seg.get();
    a.fib_sys = c.d_fib_sys.get();
    a.sys_begin = c.d_sys_begin.get();
    a.mob = c.d_mob.get();
    a.rfrig = c.d_rfrig.get();
    a.cut = c.d_cut.get();
    a.diff = c.d_diff.get();
    a.jd = c.d_jd.get();
    a.je = c.d_je.get();
    a.jd0 = c.d_jd0.get();
    a.je0 = c.d_je0.get();
    a.pdiff = c.d_pdiff.get();
    a.pcr = c.d_pcr.get();
    a.pfib = c.d_pfib.get();
    a.nseg = c.nseg;
    a.Brow = c.mB.rowptr.get();
    a.Bcol = c.mB.col.get();
    a.Bval = c.mB.val.get();
    a.Buse = c.mB.used ? 1 : 0;
    a.Crow = c.mC.rowptr.get();
    a.Ccol = c.mC.col.get();
    a.Cval = c.mC.val.get();
    a.Cuse = c.mC.used ? 1 : 0;
    a.dt = c.prm.time_step;
    return a;
}

inline SolidArgs make_solid_args(const Context& c)
{
    SolidArgs s;
    std::memset(&s, 0, sizeof(s));
    s.nsol = c.nsol;
    s.p0 = c.fib_pts;
    s.nfib = c.nfib;
    s.np = c.d_sol_np.get();
    s.idx = c.d_sol_idx.get();
    s.rad = c.d_sol_rad.get();
    s.geo = c.d_sol_geo.get();
    s.W = c.d_sol_W.get();
    s.useW = c.d_sol_useW.get();
    s.pos = c.d_pts.get();
    return s;
}

}  // namespace cymg
