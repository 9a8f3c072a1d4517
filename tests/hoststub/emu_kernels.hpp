// Host emulations of the device library's kernels, for the CPU stand-in runtime (hip/hip_runtime.h beside this
// file; test infrastructure).  They keep the control records' bookkeeping that the host code reads — k_begin's
// stop_at and the iteration count a control launch advances while an instance is live — and otherwise only read
// the first and last element that each work item addresses in its instance's arena, so AddressSanitizer reports a
// work list or descriptor table that points at freed or foreign memory.  No mechanics are computed: the kinetic
// energy and residual read 1 (NaN when node 0 of the instance holds a NaN, so a poisoned request retires as
// non-finite).
#pragma once
#include <cstdlib>

namespace wsd {
inline volatile double emu_sink = 0;
inline bool emu_live(const Ctl& c) { return !c.done && c.iter < c.stop_at; }
inline void emu_check(bool ok) { if (!ok) std::abort(); }

template <typename Real>
void k_begin(Ctl* ctls, const int32_t* sel, const int32_t* counts, int32_t nsel, int32_t n_iter, int32_t force) {
  for (int t = 0; t < nsel; t++) {
    Ctl* c = ctls + sel[t];
    c->stop_at = c->iter + (counts ? counts[t] : n_iter);
    if (force) c->done = 0;
  }
}
template <typename Real, int LANES, bool ONE>
void k_cross(const Inst<Real>* in, const CrossItem* it, int32_t n, const Ctl* ctls, int32_t) {
  for (int i = 0; i < n; i++) {
    const CrossItem& w = it[i];
    const Inst<Real>& d = in[w.inst];
    if (!emu_live(ctls[w.inst]) || w.count <= 0) continue;
    emu_check(w.c0 >= 0 && w.c0 + w.count <= d.K);
    emu_sink = emu_sink + d.u[2 * w.c0] + d.u[2 * (w.c0 + w.count) - 1] + double(d.fn[w.c0 + w.count - 1]);
  }
}
template <typename Real, int TILE, bool RB, bool ONE, int WPE, bool FUSED>
void k_strand(const Inst<Real>* in, const TileItem* it, int32_t n, Ctl* ctls) {
  for (int i = 0; i < n; i++) {
    const TileItem& t = it[i];
    const Inst<Real>& d = in[t.inst];
    if (!emu_live(ctls[t.inst])) continue;
    const int a = t.noff + t.k0, b = t.noff + t.k1 - 1;
    emu_check(a >= 0 && a <= b && b < d.n);
    emu_sink = emu_sink + d.X[0][3 * a] + d.X[1][3 * b + 2] + double(d.vx[a]) + double(d.vz[b]);
  }
}
template <typename Real, int TILE> void k_rigid(const Inst<Real>*, const TileItem*, int32_t, const Ctl*) {}
template <typename Real, bool ONE>
void k_ctrl(const Inst<Real>* in, const CtrlItem* it, int32_t n, Ctl* ctls) {
  for (int i = 0; i < n; i++) {
    const CtrlItem& w = it[i];
    if (w.chunk != 0) continue;
    Ctl& c = ctls[w.inst];
    const Inst<Real>& d = in[w.inst];
    if (!emu_live(c)) continue;
    emu_sink = emu_sink + d.X[0][3 * d.n - 1] + d.X[1][3 * d.n - 1];
    c.iter++;
    c.ke = 1.0 + 0.0 * (d.X[0][0] + d.X[1][0]);
    c.resid = c.ke;
  }
}
template <typename Real> void k_al_preview(const Inst<Real>*, int32_t, double*) {}
template <typename Real> void k_set_xq(const Inst<Real>*, int32_t, const double*, const double*) {}
template <typename Real, int TILE> void k_outputs(const Inst<Real>*, const TileItem*, int32_t, OutPtrs) {}
template <typename Real, int TILE>
void k_evaluate(const Inst<Real>*, const TileItem*, int32_t, double*, double*, double*, const int64_t*, const int64_t*) {}
#ifdef HOSTSTUB_DEFINE_COPY
void k_copy(const uint4*, uint4*, int64_t) {}
#endif
}  // namespace wsd
