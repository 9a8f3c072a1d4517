"""Results of a relaxed instance, as the reference's app reports them: buildResults (web/js/solver_cpu.js:878-922),
the results JSON export (resultsJSON, web/js/app.js:360-390) and the knots / segments tables of the reference's
reports (examples/scenes/report-*/knots.csv, segments.csv), over a downloaded state.

    r = build_results(model, state)        # state: a StateBuffers filled by GpuSolver.download (or the model itself)
    doc = results_json(model, r)           # per-strand segment arrays, knot values by mesh edge id, reactions by frame
    knots_csv(model, r), segments_csv(model, r)"""
import ctypes as C
import decimal

import numpy as np

from . import _abi as A

SEG_LAYERS = ("Mx", "My", "T", "N", "Vx", "Vy")


def _arr(st, name, n, dtype=np.float64):
    p = getattr(st.st if hasattr(st, "st") else st, name)
    if not p or n == 0:
        return np.zeros(n, dtype=dtype)
    return np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True)


def segment_moments(model, state):
    """Mx, My, T per segment from the joint and clamp moments of a downloaded state (buildResults :878-891)"""
    md = model.md
    m = md.m
    M = _arr(state, "M", 3 * md.p)
    cM = _arr(state, "cM", 2 * md.C)
    Mx, My, T = np.zeros(m), np.zeros(m), np.zeros(m)
    P = A.P_d
    rc = A.model_lib().wsm_segment_moments(C.byref(md), M.ctypes.data_as(P), cM.ctypes.data_as(P), Mx.ctypes.data_as(P),
                                           My.ctypes.data_as(P), T.ctypes.data_as(P))
    if rc != 0:
        raise ValueError(A.model_lib().wsm_last_error().decode())
    return Mx, My, T


def build_results(model, state, contact_options=None):
    """per-segment Mx My T N Vx Vy, per-crossing contact layers, per-node displacement, reactions, energies
    (buildResults, solver_cpu.js:878-922).  contact_options: which default-off layers to report
    ({"seat": bool, "rot_friction": bool, "augmented": bool}; by default those the model's parameters switch on)"""
    md = model.md
    n, m, K = md.n, md.m, md.K
    st = state.st if hasattr(state, "st") else state
    Mx, My, T = segment_moments(model, state)
    V = _arr(state, "Vsh", 2 * m)
    seg = {"Mx": Mx, "My": My, "T": T, "N": _arr(state, "N", m), "Vx": V[0::2].copy(), "Vy": V[1::2].copy()}
    X, X0 = _arr(state, "X", 3 * n), _arr(state, "X0", 3 * n)
    disp = X - X0
    par = md.par
    opt = {"seat": bool(par.seat_law), "rot_friction": bool(par.rot_friction), "augmented": bool(par.augmented)}
    opt.update(contact_options or {})
    knot = {}
    if par.contacts and K:
        phi = _arr(state, "phi", 2 * K)
        knot = {"Fn": _arr(state, "Fn", K), "Ft": _arr(state, "Ft", K), "slip": _arr(state, "slip", K),
                "state": _arr(state, "kstate", K), "gap": _arr(state, "gap", K),
                "ctype": _arr(state, "ctype", K), "pressure": _arr(state, "pressure", K),
                "util": _arr(state, "util", K), "theta": _arr(state, "theta", K),
                "phiA": phi[0::2].copy(), "phiB": phi[1::2].copy()}
        if opt["seat"]:
            dent = _arr(state, "dent", 2 * K)
            knot["dent"] = dent[0::2] + dent[1::2] + _arr(state, "dent_e", K)
        if opt["rot_friction"]:
            knot["M_rot"] = _arr(state, "Mrot", K)
            knot["rot_slip"] = _arr(state, "rot_slip", K)
        if opt["augmented"]:
            knot["lambda"] = _arr(state, "lam", K)
    else:
        knot = {k: np.zeros(K) for k in ("Fn", "Ft", "slip", "state")}
    free = np.ctypeslib.as_array(md.node_free, shape=(3 * n,)).reshape(n, 3)
    R = _arr(state, "reactions", 3 * n).reshape(n, 3)
    # the reactions of the anchored frames, keyed by frame (strand, k) as buildResults keys them by frame key
    # (solver_cpu.js:913); a fixed node that is no frame is reported under its strand-local node index
    reactions = {}
    framed = set()
    for s, k, i in model.frames():
        if not free[i].all():
            reactions[(int(s), int(k))] = [float(R[i, 0]), float(R[i, 1]), float(R[i, 2]), 0.0, 0.0, 0.0]
            framed.add(int(i))
    noff = np.ctypeslib.as_array(md.node_offset, shape=(md.S + 1,))
    for i in np.flatnonzero(~free.all(axis=1)):
        if int(i) in framed:
            continue
        s = int(np.searchsorted(noff, i, side="right") - 1)
        reactions[("node", s, int(i - noff[s]))] = [float(R[i, 0]), float(R[i, 1]), float(R[i, 2]), 0.0, 0.0, 0.0]
    # the ground (Floor.F / state and Floor.summary, solver_cpu.js:629-633, 914-915): None without one
    floor = None
    ef = st.e_floor if par.floor_on else 0.0
    if par.floor_on:
        FG = _arr(state, "floor_F", 3 * n)
        GS = _arr(state, "floor_state", n, np.uint8)
        on = GS != 0
        floor = {"z": float(par.floor_z), "mu": float(par.floor_mu), "F": FG, "state": GS, "nodes": int(on.sum()),
                 "slipping": int((GS == 2).sum()), "sum": [float(v) for v in _seq_sum(FG.reshape(n, 3)[on])],
                 "FnMax": float(st.fn_floor)}
    energies = {"kinetic": st.ke, "bending": st.e_bend, "torsion": st.e_twist, "stretch": st.e_stretch,
                "shear": st.e_shear, "contact": st.e_contact, "floor": ef,
                "beam": st.e_bend + st.e_twist + st.e_stretch + st.e_shear}
    energies["strain"] = energies["beam"] + st.e_contact + ef
    return {"iterations": st.iteration, "converged": bool(st.converged), "residual": st.residual,
            "segment": seg, "knot": knot, "disp": disp, "reactions": reactions, "floor": floor, "energies": energies,
            "rod": "cosserat"}


def _seq_sum(rows):
    """column sums of rows added one after another, in row order (the reference's summary loop)"""
    acc = [0.0, 0.0, 0.0]
    for r in rows:
        for k in range(3):
            acc[k] += float(r[k])
    return acc


def _fixed4(v):
    """a number as JavaScript's +v.toFixed(4) gives it: the nearest multiple of 1e-4 to the exact binary value,
    exact ties away from zero (toFixed rounds |v| and keeps the sign)"""
    d = decimal.Decimal(float(v))
    return float(d.quantize(decimal.Decimal("0.0001"), rounding=decimal.ROUND_HALF_UP))


def results_json(model, r, state=None):
    """the export document (resultsJSON, app.js:360-390): per-strand segment arrays, knot values by the knot's mesh
    edge id, per-strand node displacements, reactions by "strand,k" of the anchored frame, and with a ground its plane
    and the nodes resting on it (positions from `state`, else the model's own arrays)"""
    md = model.md
    soff = np.ctypeslib.as_array(md.seg_offset, shape=(md.S + 1,))
    noff = np.ctypeslib.as_array(md.node_offset, shape=(md.S + 1,))
    per_strand = lambda a: [a[soff[s]:soff[s + 1]].tolist() for s in range(md.S)]
    seg = {k: per_strand(r["segment"][k]) for k in SEG_LAYERS if k in r["segment"]}
    edges = model.knot_edges if md.K else np.zeros(0, dtype=np.int32)
    knot = {k: {str(int(edges[c])): float(v[c]) for c in range(len(v))} for k, v in r["knot"].items()}
    d = r["disp"].reshape(-1, 3)
    disp = [d[noff[s]:noff[s + 1]].tolist() for s in range(md.S)]
    reactions = {",".join(str(x) for x in key): v for key, v in r["reactions"].items()}
    doc = {"iterations": int(r["iterations"]), "converged": r["converged"], "residual": float(r["residual"]),
           "energies": r["energies"], "segment": seg, "knot": knot, "node": {"disp": disp}, "reactions": reactions}
    fl = r.get("floor")
    if fl is not None:
        # the plane and its force on every node resting on it (app.js:377-382)
        X = _arr(state if state is not None else model, "X", 3 * md.n)
        contacts = []
        for i in np.flatnonzero(fl["state"]):
            s = int(np.searchsorted(noff, i, side="right") - 1)
            contacts.append({"strand": s, "node": int(i - noff[s]), "point": [_fixed4(X[3 * i + k]) for k in range(3)],
                             "F": [float(fl["F"][3 * i + k]) for k in range(3)], "sliding": bool(fl["state"][i] == 2)})
        doc["floor"] = {"z": fl["z"], "mu": fl["mu"], "nodes": fl["nodes"], "sliding": fl["slipping"],
                        "resultant": fl["sum"], "contacts": contacts}
    return doc


def _fmt(x):
    return repr(float(x))


def knots_csv(model, r):
    """the knots table of the reference's reports (report-*/knots.csv): one row per crossing in knot-edge order,
    knot_edge,strand_a,k_a,strand_b,k_b,over,Fn,Ft,gap,slip,state"""
    kn = model.knots()
    lines = ["knot_edge,strand_a,k_a,strand_b,k_b,over,Fn,Ft,gap,slip,state"]
    k = r["knot"]
    for c in range(len(kn)):
        ids = ",".join(str(int(v)) for v in kn[c])
        vals = ",".join(_fmt(k[name][c]) for name in ("Fn", "Ft", "gap", "slip", "state"))
        lines.append(f"{ids},{vals}")
    return "\n".join(lines) + "\n"


def segments_csv(model, r):
    """the segments table of the reference's reports (report-*/segments.csv): one row per segment, strand by strand,
    strand,segment,Mx,My,T,N"""
    md = model.md
    soff = np.ctypeslib.as_array(md.seg_offset, shape=(md.S + 1,))
    seg = r["segment"]
    lines = ["strand,segment,Mx,My,T,N"]
    for s in range(md.S):
        for j in range(soff[s], soff[s + 1]):
            lines.append(f"{s},{j - soff[s]}," + ",".join(_fmt(seg[name][j]) for name in ("Mx", "My", "T", "N")))
    return "\n".join(lines) + "\n"
