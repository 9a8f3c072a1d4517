"""Closed-form pins of the oracle's crossing-contact laws that no reference fixture covers: the W1 classification
(_classify, solver_cpu.js:231-271), the smoothed support (_support, :274-282), the W3 seat friction (_seatFriction /
_walls, :285-315), the W2 rotational friction (_rotationalFriction, :347-379) and the augmented-Lagrangian multiplier
update (alValue / updateMultipliers, :480-506).

Each law is evaluated by the oracle on constructed geometry and held to what geometry and statics fix without its
own formula: the overlap of two crossing strips by polygon clipping, the support of a rectangle by its corners, the
Winkler wedge of an edge pressed into a face, the inclined-plane friction law, the friction radius of a uniformly
pressed disk and of a line contact, and the penetration an Uzawa iteration drives to zero.  The two sides of the
probed crossing get different sections and materials, so a transposed side index fails."""
import math

import numpy as np
import pytest

import cases
from oracle.bind import ContactProbe, Oracle, al_value

WA, HA, WB, HB = 60.0, 24.0, 44.0, 18.0          # section width / height of side A and side B
E90A, E90B, FCA, FCB, HSA, HSB = 700.0, 950.0, 6.0, 9.5, 28.0, 35.0


def _probe():
    m = cases.crossing()
    for name, vals in (("k_wk", (WA, WB)), ("k_hk", (HA, HB)), ("k_E90", (E90A, E90B)), ("k_fc90", (FCA, FCB)),
                       ("k_Hs", (HSA, HSB))):
        m.arr(name, 2)[:] = vals
    return ContactProbe(m)


def frame(angle, roll=0.0):
    """a lath along (cos a, sin a, 0) lying flat (height axis y = +z), rolled by `roll` about its tangent: rows x
    (width axis), y (height axis), t"""
    t = np.array([math.cos(angle), math.sin(angle), 0.0])
    y = np.array([0.0, 0.0, 1.0])
    x = np.cross(y, t)
    xr = math.cos(roll) * x + math.sin(roll) * y
    yr = -math.sin(roll) * x + math.cos(roll) * y
    return np.array([xr, yr, t])


Z = np.array([0.0, 0.0, 1.0])


# ---------------------------------------------------------------------------------------------- polygon clipping
def _clip(poly, a, b):
    """Sutherland-Hodgman: the part of polygon `poly` left of the directed line a→b"""
    def inside(p):
        return (b[0] - a[0]) * (p[1] - a[1]) - (b[1] - a[1]) * (p[0] - a[0]) >= 0
    out = []
    for i in range(len(poly)):
        p, q = poly[i], poly[(i + 1) % len(poly)]
        if inside(q):
            if not inside(p):
                out.append(_cross_pt(p, q, a, b))
            out.append(q)
        elif inside(p):
            out.append(_cross_pt(p, q, a, b))
    return out


def _cross_pt(p, q, a, b):
    d1, d2 = np.subtract(q, p), np.subtract(b, a)
    t = ((a[0] - p[0]) * d2[1] - (a[1] - p[1]) * d2[0]) / (d1[0] * d2[1] - d1[1] * d2[0])
    return (p[0] + t * d1[0], p[1] + t * d1[1])


def _strip(angle, width, length=1e4):
    t = np.array([math.cos(angle), math.sin(angle)])
    nrm = np.array([-t[1], t[0]])
    c = [length * t + 0.5 * width * nrm, -length * t + 0.5 * width * nrm, -length * t - 0.5 * width * nrm,
         length * t - 0.5 * width * nrm]
    pts = [tuple(p) for p in c]
    # counter-clockwise order
    area = sum(pts[i][0] * pts[(i + 1) % 4][1] - pts[(i + 1) % 4][0] * pts[i][1] for i in range(4))
    return pts if area > 0 else pts[::-1]


def overlap_area(angle_a, width_a, angle_b, width_b):
    P = _strip(angle_a, width_a)
    Qp = _strip(angle_b, width_b)
    for i in range(4):
        P = _clip(P, Qp[i], Qp[(i + 1) % 4])
    return 0.5 * abs(sum(P[i][0] * P[(i + 1) % len(P)][1] - P[(i + 1) % len(P)][0] * P[i][1] for i in range(len(P))))


# ---------------------------------------------------------------------------------------------- W1 classification
@pytest.mark.parametrize("cross_deg", [90.0, 60.0, 35.0, 120.0])
def test_face_on_face_area_is_the_strip_overlap(cross_deg):
    """two flat laths: face-on-face, θ the acute crossing angle, A_nom the overlap of the two strips (polygon
    clipping), pressure = Fn / A_nom and utilisation = pressure / min H"""
    p = _probe()
    th = math.radians(cross_deg)
    fn = 123.0
    r = p.classify(frame(0.0), frame(th), Z, fn)
    assert r["ctype"] == 0 and r["face_side"] == -1
    assert r["theta"] == pytest.approx(min(th, math.pi - th), rel=1e-12)
    area = overlap_area(0.0, WA, th, WB)
    assert r["Anom"] == pytest.approx(area, rel=1e-9)
    assert r["pressure"] == pytest.approx(fn / area, rel=1e-9)
    assert r["util"] == pytest.approx(fn / area / min(HSA, HSB), rel=1e-9)
    assert r["phi"] == (0.0, 0.0)


def test_lath_standing_on_its_narrow_face_uses_its_height():
    """A rolled 90° (on edge) lies on its narrow face: still face-on-face, the overlap strip of A is HA wide"""
    p = _probe()
    th = math.radians(70.0)
    r = p.classify(frame(0.0, math.pi / 2), frame(th), Z, 50.0)
    assert r["ctype"] == 0 and r["cface"] == (pytest.approx(HA), pytest.approx(WB))
    assert r["Anom"] == pytest.approx(overlap_area(0.0, HA, th, WB), rel=1e-9)


@pytest.mark.parametrize("roll_deg", [12.0, 30.0, 45.0, 63.0])
def test_roll_angle_and_contact_type(roll_deg):
    """a section rolled off the contact plane by more than the smoothing width touches with an edge; φ is the roll
    folded into [0°, 45°]; one rolled side gives edge-on-face with the flat side as the face, two give edge-on-edge"""
    p = _probe()
    al = math.radians(roll_deg)
    r = p.classify(frame(0.0, al), frame(math.radians(80.0)), Z, 40.0)
    assert r["ctype"] == 1 and r["face_side"] == 1
    assert r["phi"][0] == pytest.approx(min(al, math.pi / 2 - al), abs=1e-12) and r["phi"][1] == 0.0
    r = p.classify(frame(0.0), frame(math.radians(80.0), -al), Z, 40.0)
    assert r["ctype"] == 1 and r["face_side"] == 0
    r = p.classify(frame(0.0, al), frame(math.radians(80.0), al), Z, 40.0)
    assert r["ctype"] == 2 and r["face_side"] == -1


@pytest.mark.parametrize("fn", [2.0, 15.0, 60.0])
def test_edge_on_face_is_a_winkler_wedge(fn):
    """the rolled edge of A pressed into the flat face of B: B's face is a Winkler foundation of modulus
    κ = E90_B / h_B; the edge's flank leaves the face at the roll angle φ, so the pressure rises linearly across the
    contact width b, p(x) = κ φ x, along the contact line ℓ = w_B / sinθ (the chord of B's face).  Statics:
    Fn = ½ κ φ b² ℓ; the nominal area b ℓ, the peak pressure κ φ b, utilisation against the crushing strength of the
    edge side (A)"""
    p = _probe()
    th, al = math.radians(75.0), math.radians(20.0)
    r = p.classify(frame(0.0, al), frame(th), Z, fn)
    assert r["ctype"] == 1 and r["face_side"] == 1
    ell = WB / math.sin(th)
    kap = E90B / HB
    b = r["Anom"] / ell
    assert b < WA                                        # the wedge has not flattened across A's face
    assert 0.5 * kap * al * b * b * ell == pytest.approx(fn, rel=1e-12)
    assert r["pressure"] == pytest.approx(kap * al * b, rel=1e-12)
    assert r["util"] == pytest.approx(r["pressure"] / FCA, rel=1e-12)


def test_edge_on_face_flattened_wedge_takes_the_full_overlap():
    """a load large enough to press the wedge flat across A's face: the full parallelogram carries it"""
    p = _probe()
    th, al = math.radians(75.0), math.radians(20.0)
    fn = 0.5 * (E90B / HB) * al * (2 * WA) ** 2 * (WB / math.sin(th))     # a wedge twice as wide as A's face
    r = p.classify(frame(0.0, al), frame(th), Z, fn)
    full = overlap_area(0.0, WA, th, WB)
    assert r["Anom"] == pytest.approx(full, rel=1e-9)
    assert r["pressure"] == pytest.approx(fn / full, rel=1e-9)
    assert r["util"] == pytest.approx(fn / full / FCA, rel=1e-9)


def test_edge_on_edge_crushes_at_the_weaker_strength():
    """two edges: the contact crushes until it carries Fn at the weaker crushing strength"""
    p = _probe()
    r = p.classify(frame(0.0, 0.4), frame(1.2, 0.7), Z, 33.0)
    assert r["ctype"] == 2
    assert r["Anom"] == pytest.approx(33.0 / min(FCA, FCB), rel=1e-12)
    assert r["pressure"] == min(FCA, FCB) and r["util"] == 1.0
    r = p.classify(frame(0.0, 0.4), frame(1.2, 0.7), Z, 0.0)
    assert r["pressure"] == 0.0 and r["util"] == 0.0


# ---------------------------------------------------------------------------------------------- support
@pytest.mark.parametrize("roll_deg", [0.0, 90.0, 17.0, 45.0, 71.0, 180.0])
@pytest.mark.parametrize("side,sign", [(0, 1.0), (1, -1.0)])
def test_support_is_the_section_extent(roll_deg, side, sign):
    """the half-extent of a w × h rectangle towards sign·n is the largest projection of its corners; the support
    point is the face centre of an aligned section and the extreme corner of a rolled one (within the tanh
    smoothing, e^(-2 |n·axis| / 0.03))"""
    p = _probe()
    w, h = (WA, HA) if side == 0 else (WB, HB)
    R = frame(0.3, math.radians(roll_deg))
    hs, r = p.support(R, Z, sign, side)
    d = sign * Z
    corners = [sx * 0.5 * w * R[0] + sy * 0.5 * h * R[1] for sx in (-1, 1) for sy in (-1, 1)]
    proj = [float(c @ d) for c in corners]
    assert hs == pytest.approx(max(proj), rel=1e-12, abs=1e-12)
    nx, ny = float(R[0] @ d), float(R[1] @ d)
    if abs(nx) < 1e-12:                 # aligned: the centre of the face
        assert np.allclose(r, math.copysign(0.5 * h, ny) * R[1], atol=1e-9)
    elif abs(ny) < 1e-12:
        assert np.allclose(r, math.copysign(0.5 * w, nx) * R[0], atol=1e-9)
    else:
        best = corners[int(np.argmax(proj))]
        tol = 0.5 * max(w, h) * 2 * math.exp(-2 * min(abs(nx), abs(ny)) / 0.03) + 1e-9
        assert np.abs(r - best).max() <= tol


def test_aligned_supports_sum_to_the_half_heights():
    """two flat laths: D = h_A + h_B is half of each section's height, the centre-line separation when the faces
    touch"""
    p = _probe()
    hA, _ = p.support(frame(0.0), Z, 1.0, 0)
    hB, _ = p.support(frame(1.0), Z, -1.0, 1)
    assert hA + hB == 0.5 * HA + 0.5 * HB


# ---------------------------------------------------------------------------------------------- W3 seat friction
@pytest.mark.parametrize("roll_deg", [10.0, 25.0, 40.0])
@pytest.mark.parametrize("mu", [0.1, 0.35])
def test_seat_friction_is_the_inclined_wall_law(roll_deg, mu):
    """A sliding out of the dent that B cut into it, across B (slip ⟂ t_B in the contact plane), climbs a wall of
    B's rolled section inclined by ψ (one of B's faces: tan ψ = tan α or cot α): the friction of a wedge pushed up an
    incline, μ_eff = tan(atan μ + ψ).  Sliding along t_B meets no wall: μ_eff = μ"""
    m = cases.crossing(mu=mu)
    for name, vals in (("k_wk", (WA, WB)), ("k_hk", (HA, HB))):
        m.arr(name, 2)[:] = vals
    p = ContactProbe(m)
    al = math.radians(roll_deg)
    RA, RB = frame(0.0), frame(math.pi / 2, al)
    tB = RB[2]
    across = np.cross(Z, tB)
    for d in (across, -across):
        mu_eff, tp = p.seat(RA, RB, Z, 0.7 * d, (0.3, 0.0))
        psi = math.atan(tp[0])
        assert tp[0] == pytest.approx(math.tan(al), rel=1e-9) or tp[0] == pytest.approx(1 / math.tan(al), rel=1e-9)
        law = math.tan(math.atan(mu) + psi) if mu * tp[0] < 1 else 10 * mu
        assert mu_eff == pytest.approx(min(law, 10 * mu), rel=1e-9)
    mu_eff, _ = p.seat(RA, RB, Z, 0.7 * tB, (0.3, 0.0))
    assert mu_eff == pytest.approx(mu, abs=1e-12)
    mu_eff, _ = p.seat(RA, RB, Z, 0.7 * across, (0.0, 0.0))       # no dent: plain Coulomb
    assert mu_eff == mu


def test_seat_friction_blends_with_the_slip_direction():
    """slip at angle β to the wall normal: the wall's share of the extra friction goes with |cos β|"""
    p = _probe()
    al = math.radians(30.0)
    RA, RB = frame(0.0), frame(math.pi / 2, al)
    tB = RB[2]
    across = np.cross(Z, tB)
    mu0 = 0.35
    full, tp = p.seat(RA, RB, Z, across, (0.3, 0.0))
    for beta in (0.3, 0.9, 1.3):
        d = math.cos(beta) * across + math.sin(beta) * tB
        mu_eff, _ = p.seat(RA, RB, Z, 2.0 * d, (0.3, 0.0))
        assert mu_eff == pytest.approx(mu0 + abs(math.cos(beta)) * (full - mu0), rel=1e-9)


# ---------------------------------------------------------------------------------------------- W2
def _rot_case(ctype):
    """frames and the expected friction radius r_eq of a face (uniform disk of area A_nom: ⅔ √(A/π)), an
    edge-on-face line contact (ℓ / 4, ℓ = w_face / sinθ) or an edge-on-edge contact (a disk again)"""
    th = math.radians(70.0)
    if ctype == 0:
        return frame(0.0), frame(th), th
    if ctype == 1:
        return frame(0.0, 0.35), frame(th), th
    return frame(0.0, 0.35), frame(th, 0.5), th


@pytest.mark.parametrize("ctype", [0, 1, 2])
def test_rotational_friction_sticks_then_sits_on_the_coulomb_moment(ctype):
    """an imposed twist δ about n: below the Coulomb moment μ Fn r_eq the spring k_t r_eq² δ sticks, above it the
    moment is returned onto μ Fn r_eq exactly (slip rate 1) and the reference angle moves by the excess"""
    p = _probe()
    RA, RB, th = _rot_case(ctype)
    fn, kt, mu = 80.0, 1e4, 0.35
    c = p.classify(RA, RB, Z, fn)
    assert c["ctype"] == ctype
    if ctype == 1:
        ell = WB / math.sin(th)                       # B is the face side: the contact line spans B's width
        req = ell / 4
        rel = 1e-9                                    # the twist is formed from the frames' rounded angles
    else:
        req = 2.0 / 3.0 * math.sqrt(c["Anom"] / math.pi)
        rel = 0.011                                   # the reference's 0.38 against ⅔/√π = 0.3761
    mlim = mu * fn * req
    kr = kt * req * req
    small = 0.2 * mlim / kr
    r = p.rot(RA, RB, Z, fn, th - small, True, kt)
    assert r["M"] == pytest.approx(kr * small, rel=2 * rel) and r["rot_slip"] == 0.0
    big = 3.0 * mlim / kr
    assert big < 0.05                                 # within the per-iteration advance clip
    r = p.rot(RA, RB, Z, fn, th - big, True, kt)
    assert r["M"] == pytest.approx(mlim, rel=rel)
    # the stick point moved by the excess rotation: the remaining twist carries the Coulomb moment
    rem = th - r["theta0"]
    assert r["rot_slip"] == pytest.approx(big - rem, rel=1e-9)
    assert kr * rem == pytest.approx(r["M"], rel=2 * rel)
    r = p.rot(RA, RB, Z, fn, th - big, False, kt)     # not in contact: no moment
    assert r["M"] == 0.0


def test_rotational_friction_wraps_the_twist_and_sets_the_reference_once():
    p = _probe()
    RA, RB, th = _rot_case(0)
    a = p.rot(RA, RB, Z, 80.0, th - 1e-4, True, 1e4)
    b = p.rot(RA, RB, Z, 80.0, th - 1e-4 - 2 * math.pi, True, 1e4)
    assert b["M"] == pytest.approx(a["M"], rel=1e-9)
    r = p.rot(RA, RB, Z, 80.0, float("nan"), True, 1e4)     # unset: taken from the current angle, no moment
    assert r["theta0"] == pytest.approx(th, rel=1e-12) and r["M"] == 0.0


# ---------------------------------------------------------------------------------------------- AL
def test_al_value_safeguards():
    """λ ← max(0, λ + k_n pen̄), bilateral when locked, 0 at a pinned or disabled crossing, |λ| ≤ cap"""
    assert al_value(10.0, 1e-3, 1e4, 0, 1, 0, 1e9) == pytest.approx(20.0)
    assert al_value(10.0, -5e-3, 1e4, 0, 1, 0, 1e9) == 0.0
    assert al_value(10.0, -5e-3, 1e4, 1, 1, 0, 1e9) == pytest.approx(-40.0)
    assert al_value(10.0, 1.0, 1e4, 0, 1, 0, 25.0) == 25.0
    assert al_value(10.0, -1.0, 1e4, 1, 1, 0, 25.0) == -25.0
    assert al_value(10.0, 1e-3, 1e4, 0, 0, 0, 1e9) == 0.0
    assert al_value(10.0, 1e-3, 1e4, 0, 1, 1, 1e9) == 0.0


def _pressed(**kw):
    m = cases.crossing(ramp_iter=0, tol=1e-9, **kw)
    o = Oracle(m)
    o.run(200000, 5000)
    return m, o, o.state()


# The pressed crossing: 50 N down on B at the knot; A and B are the same beam (same section, span and end supports
# in z), so the knot carries a two-spring split.  With beam stiffness k at mid-span and contact penetration pen:
# Fn = k δA,  50 − Fn = k δB,  δB − δA = pen  ⇒  Fn = (50 − k pen) / 2.  A penalty contact has pen = Fn / k_n, so
# Fn_penalty = 50 / (2 + k / k_n), which measures k / k_n; a rigid contact (pen = 0) carries exactly half the load.
KN = 1e4


@pytest.mark.slow
def test_augmented_lagrangian_drives_the_penetration_to_zero():
    """the augmented form's Uzawa updates drive the penetration to zero: the knot carries exactly half the load, all
    of it in the multiplier, and the gap opens by the penalty form's depth"""
    mp, op, sp = _pressed()
    ma, oa, sa = _pressed(augmented=1, al_every=200)
    assert op.converged and oa.converged
    pen_p = sp["Fn"][0] / KN
    k_over_kn = 50.0 / sp["Fn"][0] - 2.0
    assert 0 < k_over_kn < 0.05                          # the penalty form's own split: slightly under half
    assert sa["Fn"][0] == pytest.approx(25.0, rel=1e-3)
    assert sa["lam"][0] == pytest.approx(sa["Fn"][0], rel=1e-3)
    pen_a = (sa["Fn"][0] - sa["lam"][0]) / KN
    assert abs(pen_a) <= 1e-2 * pen_p
    assert sa["gap"][0] - sp["gap"][0] == pytest.approx(pen_p - pen_a, rel=2e-2)


@pytest.mark.slow
def test_augmented_lagrangian_multiplier_stops_at_the_crushing_cap():
    """a crushing strength so low that fc90·wA·wB/sinθ is below the knot's share: λ stays at that cap and the
    penetration carries the rest, pen = (Fn − cap) / k_n, with the split of the two-spring model,
    Fn = (50 + (k/k_n) cap) / (2 + k/k_n), k/k_n measured by the penalty form"""
    fc = 0.004
    cap = fc * 60.0 * 60.0
    assert cap < 25.0
    _, op, sp = _pressed(fc90=fc)
    k_over_kn = 50.0 / sp["Fn"][0] - 2.0
    m, o, s = _pressed(augmented=1, al_every=200, fc90=fc)
    assert o.converged
    assert s["lam"][0] == pytest.approx(cap, rel=1e-12)
    assert s["Fn"][0] == pytest.approx((50.0 + k_over_kn * cap) / (2.0 + k_over_kn), rel=1e-3)
