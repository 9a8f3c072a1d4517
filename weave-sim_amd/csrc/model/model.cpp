// Host model builder: scene + mesh → kagome topology → solver state → goals (f64), as the reference's host side
// builds a CpuSolver before GpuSolver uploads it.  Each function names the reference code whose behaviour it follows.
// Build: g++ -O2 -std=c++17 -ffp-contract=off -fPIC -shared (no fast-math: the arithmetic follows the reference's
// f64 formulas operation by operation).
#include "weave_model.h"

#include <algorithm>
#include <array>
#include <tuple>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <set>
#include <memory>
#include <numeric>
#include <queue>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

using V3 = std::array<double, 3>;
inline V3 sub3(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 add3(const V3& a, const V3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 scl3(const V3& a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
inline double dot3(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 cross3(const V3& a, const V3& b) { return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}; }
inline double norm3(const V3& a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
// topology.js:28 normalize3 (a zero vector stays zero)
inline V3 normalize3(const V3& v, double eps = 1e-12) { double n = norm3(v); if (n < eps) n = 1.0; return {v[0] / n, v[1] / n, v[2] / n}; }

// ------------------------------------------------------------------------------------------------ rng.js:4-27
// xoshiro128** seeded by splitmix32 (the generator of the reference's seeded noise)
struct Rng {
  uint32_t s[4];
  explicit Rng(uint32_t seed) {
    uint32_t a = seed;
    auto sm = [&a]() {
      a = a + 0x9e3779b9u;
      uint32_t t = a ^ (a >> 16); t = t * 0x21f0aaadu; t ^= t >> 15; t = t * 0x735a2d97u;
      return t ^ (t >> 15);
    };
    for (auto& w : s) w = sm();
  }
  static uint32_t rotl(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }
  uint32_t u32() {
    uint32_t r = rotl(s[1] * 5u, 7) * 9u, t = s[1] << 9;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 11);
    return r;
  }
  double uniform() { double hi = double(u32() >> 5), lo = double(u32() >> 6); return (hi * 67108864.0 + lo) / 9007199254740992.0; }
};

// ------------------------------------------------------------------------------------------------ quat.js
void toFrame(const double* q, double* R) {  // quat.js:122-130
  double w = q[0], x = q[1], y = q[2], z = q[3];
  double n = std::sqrt(w * w + x * x + y * y + z * z); if (n == 0) n = 1.0;
  w /= n; x /= n; y /= n; z /= n;
  R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y + z * w); R[2] = 2 * (x * z - y * w);
  R[3] = 2 * (x * y - z * w); R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z + x * w);
  R[6] = 2 * (x * z + y * w); R[7] = 2 * (y * z - x * w); R[8] = 1 - 2 * (x * x + y * y);
}
void qnormalize(double* q) {  // quat.js:107-111
  double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (n == 0) n = 1.0;
  q[0] /= n; q[1] /= n; q[2] /= n; q[3] /= n;
}
void fromFrame(const double* F, double* out) {  // quat.js:132-142 (setup version: strict comparisons)
  double m00 = F[0], m10 = F[1], m20 = F[2], m01 = F[3], m11 = F[4], m21 = F[5], m02 = F[6], m12 = F[7], m22 = F[8];
  double t = m00 + m11 + m22;
  if (t > 0) { double s = std::sqrt(t + 1.0) * 2; out[0] = 0.25 * s; out[1] = (m21 - m12) / s; out[2] = (m02 - m20) / s; out[3] = (m10 - m01) / s; }
  else if (m00 > m11 && m00 > m22) { double s = std::sqrt(1.0 + m00 - m11 - m22) * 2; out[0] = (m21 - m12) / s; out[1] = 0.25 * s; out[2] = (m01 + m10) / s; out[3] = (m02 + m20) / s; }
  else if (m11 > m22) { double s = std::sqrt(1.0 + m11 - m00 - m22) * 2; out[0] = (m02 - m20) / s; out[1] = (m01 + m10) / s; out[2] = 0.25 * s; out[3] = (m12 + m21) / s; }
  else { double s = std::sqrt(1.0 + m22 - m00 - m11) * 2; out[0] = (m10 - m01) / s; out[1] = (m02 + m20) / s; out[2] = (m12 + m21) / s; out[3] = 0.25 * s; }
  qnormalize(out);
}
// rod_cosserat.js:33-45: Ψ = log(q_a* ⊗ q_b), returns θ
double relRotation(const double* qa_, const double* qb, double* out) {
  double aw = qa_[0], ax = -qa_[1], ay = -qa_[2], az = -qa_[3];
  double bw = qb[0], bx = qb[1], by = qb[2], bz = qb[3];
  double w = aw * bw - ax * bx - ay * by - az * bz;
  double x = aw * bx + ax * bw + ay * bz - az * by;
  double y = aw * by - ax * bz + ay * bw + az * bx;
  double z = aw * bz + ax * by - ay * bx + az * bw;
  if (w < 0) { w = -w; x = -x; y = -y; z = -z; }
  double s = std::sqrt(x * x + y * y + z * z), th = 2 * std::atan2(s, w);
  double f = s > 1e-12 ? th / s : 2 / std::max(w, 1e-12);
  out[0] = f * x; out[1] = f * y; out[2] = f * z;
  return th;
}

// topology.js:37-44: R with Z = tangent, Y = hint ⟂ Z, X = Y × Z
std::array<double, 9> frameFromTangentNormal(const V3& t, const V3& hint) {
  V3 z = normalize3(t);
  V3 y = sub3(hint, scl3(z, dot3(hint, z)));
  if (norm3(y) < 1e-9) { V3 a = std::fabs(z[0]) < 0.9 ? V3{1, 0, 0} : V3{0, 1, 0}; y = sub3(a, scl3(z, dot3(a, z))); }
  y = normalize3(y);
  V3 x = cross3(y, z);
  return {x[0], x[1], x[2], y[0], y[1], y[2], z[0], z[1], z[2]};
}

// ------------------------------------------------------------------------------------------------ mesh.js
struct Mesh {
  std::vector<double> V; std::vector<int32_t> F;
  int nV = 0, nF = 0, nE = 0;
  std::vector<int32_t> E, FE, EF;
  std::vector<double> faceN, edgeN, edgeMid;
  std::vector<uint8_t> isBnd;
  int otherFace(int e, int f) const { int a = EF[2 * e], b = EF[2 * e + 1]; return a == f ? b : a; }
  bool edgeHas(int e, int v) const { return E[2 * e] == v || E[2 * e + 1] == v; }
  int faceEdgeContaining(int f, int v, int exclude) const {  // mesh.js:42-49
    for (int i = 0; i < 3; i++) { int e = FE[3 * f + i]; if (e == exclude) continue; if (edgeHas(e, v)) return e; }
    throw std::runtime_error("face has no edge containing the vertex");
  }
};

double roundHalfEven(double x) {  // mesh.js:104 (np.round)
  double r = std::floor(x + 0.5);
  return std::fabs(std::fmod(x, 1.0)) == 0.5 ? 2 * std::floor(x / 2 + 0.5) : r;
}

// mesh.js:108-133: weld duplicated vertices (quantised at tol × model size, lexicographic unique rows)
void weld(std::vector<double>& V, std::vector<int32_t>& F, double tol = 1e-9) {
  size_t n = V.size() / 3;
  if (!n) return;
  double ptp = 0;
  for (int c = 0; c < 3; c++) { double lo = INFINITY, hi = -INFINITY; for (size_t i = 0; i < n; i++) { double v = V[3 * i + c]; lo = std::min(lo, v); hi = std::max(hi, v); } ptp = std::max(ptp, hi - lo); }
  double scale = std::max(ptp, 1.0), q = scale * tol;
  std::vector<double> key(3 * n);
  for (size_t i = 0; i < 3 * n; i++) key[i] = roundHalfEven(V[i] / q);
  std::vector<int> order(n); std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    for (int c = 0; c < 3; c++) if (key[3 * a + c] != key[3 * b + c]) return key[3 * a + c] < key[3 * b + c];
    return a < b;
  });
  std::vector<int32_t> inv(n); std::vector<int> first; int u = -1;
  for (size_t r = 0; r < n; r++) {
    int i = order[r];
    if (r == 0 || key[3 * i] != key[3 * order[r - 1]] || key[3 * i + 1] != key[3 * order[r - 1] + 1] || key[3 * i + 2] != key[3 * order[r - 1] + 2]) { u++; first.push_back(i); }
    inv[i] = u;
  }
  std::vector<double> Vw(3 * first.size());
  for (size_t k = 0; k < first.size(); k++) for (int c = 0; c < 3; c++) Vw[3 * k + c] = V[3 * first[k] + c];
  std::vector<int32_t> Fw;
  for (size_t f = 0; f < F.size() / 3; f++) {
    int a = inv[F[3 * f]], b = inv[F[3 * f + 1]], c = inv[F[3 * f + 2]];
    if (a != b && b != c && a != c) { Fw.push_back(a); Fw.push_back(b); Fw.push_back(c); }
  }
  V.swap(Vw); F.swap(Fw);
}

// mesh.js:8-21 fromArrays with edgeTopology (:137-163), face/edge normals (:165-187)
Mesh meshFromArrays(const std::vector<double>& V, const std::vector<int32_t>& F) {
  Mesh m; m.V = V; m.F = F; m.nV = int(V.size() / 3); m.nF = int(F.size() / 3);
  int nF = m.nF;
  std::vector<std::pair<int, int>> rows(3 * nF);
  for (int i = 0; i < 3; i++) for (int f = 0; f < nF; f++) {
    int a = F[3 * f + (i + 1) % 3], b = F[3 * f + (i + 2) % 3];
    rows[i * nF + f] = a < b ? std::make_pair(a, b) : std::make_pair(b, a);
  }
  std::vector<int> order(3 * nF); std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int r, int s) { return rows[r] < rows[s] || (rows[r] == rows[s] && r < s); });
  std::vector<int32_t> inv(3 * nF); int u = -1;
  for (size_t k = 0; k < order.size(); k++) {
    int r = order[k];
    if (k == 0 || rows[r] != rows[order[k - 1]]) { u++; m.E.push_back(rows[r].first); m.E.push_back(rows[r].second); }
    inv[r] = u;
  }
  m.nE = int(m.E.size() / 2);
  m.FE.assign(3 * nF, 0);
  for (int i = 0; i < 3; i++) for (int f = 0; f < nF; f++) m.FE[3 * f + i] = inv[i * nF + f];
  m.EF.assign(2 * m.nE, -1);
  for (int r = 0; r < 3 * nF; r++) {
    int e = inv[r], f = r % nF;
    if (m.EF[2 * e] < 0) m.EF[2 * e] = f;
    else if (m.EF[2 * e + 1] < 0) m.EF[2 * e + 1] = f;
    else throw std::runtime_error("non-manifold edge shared by more than two faces");
  }
  m.faceN.assign(3 * nF, 0.0);
  for (int f = 0; f < nF; f++) {
    int i = 3 * F[3 * f], j = 3 * F[3 * f + 1], k = 3 * F[3 * f + 2];
    double ux = V[j] - V[i], uy = V[j + 1] - V[i + 1], uz = V[j + 2] - V[i + 2], vx = V[k] - V[i], vy = V[k + 1] - V[i + 1], vz = V[k + 2] - V[i + 2];
    double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
    double l = std::sqrt(nx * nx + ny * ny + nz * nz); if (l == 0) l = 1;
    m.faceN[3 * f] = nx / l; m.faceN[3 * f + 1] = ny / l; m.faceN[3 * f + 2] = nz / l;
  }
  m.edgeN.assign(3 * m.nE, 0.0);
  for (int side = 0; side < 2; side++) for (int e = 0; e < m.nE; e++) { int f = m.EF[2 * e + side]; if (f >= 0) for (int c = 0; c < 3; c++) m.edgeN[3 * e + c] += m.faceN[3 * f + c]; }
  for (int e = 0; e < m.nE; e++) {
    double l = std::sqrt(m.edgeN[3 * e] * m.edgeN[3 * e] + m.edgeN[3 * e + 1] * m.edgeN[3 * e + 1] + m.edgeN[3 * e + 2] * m.edgeN[3 * e + 2]); if (l == 0) l = 1;
    for (int c = 0; c < 3; c++) m.edgeN[3 * e + c] /= l;
  }
  m.isBnd.assign(m.nE, 0);
  for (int e = 0; e < m.nE; e++) m.isBnd[e] = (m.EF[2 * e] < 0 || m.EF[2 * e + 1] < 0) ? 1 : 0;
  m.edgeMid.assign(3 * m.nE, 0.0);
  for (int e = 0; e < m.nE; e++) for (int c = 0; c < 3; c++) m.edgeMid[3 * e + c] = 0.5 * (V[3 * m.E[2 * e] + c] + V[3 * m.E[2 * e + 1] + c]);
  return m;
}

// ------------------------------------------------------------------------------------------------ topology.js
struct Strand { std::vector<int> edges; std::vector<std::pair<int, int>> segs; bool closed = false; };
struct Knot { int edge; int as, ak, bs, bk; int overSide = 0; bool conflict = false, flipped = false, isBoundary = false; };
inline int64_t fkey(int s, int k) { return int64_t(s) * 1048576 + k; }

struct Topology {
  const Mesh* mesh = nullptr;
  int subdivisions = 4; double overhang = 0.0; bool halfHeight = true; int seedPhase = 0;
  double thickness = 30.0;
  std::vector<double> strandH;                // hybrid weave: the height of each strand's lath group; empty otherwise
  std::vector<Strand> strands;
  std::vector<int> knotOrder;                 // edge ids in insertion order (topology.js Map order)
  std::map<int, Knot> knots;                  // edge id → knot
  std::map<int64_t, int> frameToKnot;         // fkey → edge id
  std::map<int64_t, int> knotNode;            // fkey → polyline row
  std::vector<std::vector<double>> polylines; // per strand 3 N_s
  std::vector<std::vector<double>> nodeFrames;// per strand 9 N_s
  int parityConflicts = 0;
  int nStrands() const { return int(strands.size()); }
};

bool isOverFrame(const Knot& kn, int s, int k) {  // topology.js:18-24
  if (kn.overSide) return kn.bs == s && kn.bk == k;
  return kn.as == s && kn.ak == k;
}
void setOver(Knot& kn, int side) { kn.overSide = side ? 1 : 0; }

// topology.js:90-149
void traceStrands(Topology& topo) {
  const Mesh& mesh = *topo.mesh;
  const std::vector<int32_t>& E = mesh.E;
  std::vector<uint8_t> visited(mesh.nF, 0);   // bit c: corner c of the face visited
  auto corner = [&](int f, int v) { for (int c = 0; c < 3; c++) if (mesh.F[3 * f + c] == v) return c; throw std::runtime_error("vertex not on face"); };
  auto isVis = [&](int f, int v) { return (visited[f] >> corner(f, v)) & 1; };
  auto setVis = [&](int f, int v) { visited[f] |= uint8_t(1u << corner(f, v)); };
  std::vector<Strand> strands;
  auto finish = [&](std::vector<int> edges, std::vector<std::pair<int, int>> segs, bool closed) {
    if (closed && edges.size() > segs.size()) edges.resize(segs.size());
    Strand s; s.edges = std::move(edges); s.segs = std::move(segs); s.closed = closed; return s;
  };
  auto trace = [&](int face, int entry, int pivot, Strand& out) -> bool {
    std::vector<int> edges{entry}; std::vector<std::pair<int, int>> segs;
    const int f0 = face, e0 = entry, p0 = pivot;
    for (;;) {
      int q = pivot;
      int p = E[2 * entry] == q ? E[2 * entry + 1] : E[2 * entry];
      if (isVis(face, p)) {
        bool closed = !segs.empty() && face == f0 && entry == e0 && pivot == p0;
        if (segs.empty()) return false;
        out = finish(edges, segs, closed); return true;
      }
      setVis(face, p);
      segs.push_back({face, p});
      int g = mesh.faceEdgeContaining(face, p, entry);
      int nxt = mesh.otherFace(g, face);
      pivot = p; entry = g;
      if (nxt < 0) { edges.push_back(g); out = finish(edges, segs, false); return true; }
      face = nxt;
      if (face == f0 && entry == e0 && pivot == p0) { out = finish(edges, segs, true); return true; }
      edges.push_back(g);
    }
  };
  for (int e = 0; e < mesh.nE; e++) {
    if (!mesh.isBnd[e]) continue;
    int f = mesh.EF[2 * e] >= 0 ? mesh.EF[2 * e] : mesh.EF[2 * e + 1];
    for (int pv : {E[2 * e], E[2 * e + 1]}) { Strand st; if (trace(f, e, pv, st)) strands.push_back(std::move(st)); }
  }
  for (int f = 0; f < mesh.nF; f++) for (int c = 0; c < 3; c++) {
    int cv = mesh.F[3 * f + c];
    if (isVis(f, cv)) continue;
    int entry = -1;
    for (int i = 0; i < 3; i++) { int e = mesh.FE[3 * f + i]; if (mesh.edgeHas(e, cv)) { entry = e; break; } }
    int otherV = E[2 * entry] == cv ? E[2 * entry + 1] : E[2 * entry];
    Strand st; if (trace(f, entry, otherV, st)) strands.push_back(std::move(st));
  }
  topo.strands = std::move(strands);
  std::map<int, std::vector<std::pair<int, int>>> passes; std::vector<int> order;
  for (int s = 0; s < topo.nStrands(); s++) {
    const auto& ed = topo.strands[s].edges;
    for (int k = 0; k < int(ed.size()); k++) { auto it = passes.find(ed[k]); if (it == passes.end()) { order.push_back(ed[k]); passes[ed[k]] = {}; } passes[ed[k]].push_back({s, k}); }
  }
  for (int e : order) {
    const auto& fids = passes[e];
    if (fids.size() != 2) throw std::runtime_error("edge visited " + std::to_string(fids.size()) + " times; tracing invariant broken");
    Knot kn{e, fids[0].first, fids[0].second, fids[1].first, fids[1].second};
    kn.isBoundary = mesh.isBnd[e] != 0;
    topo.knots[e] = kn; topo.knotOrder.push_back(e);
    topo.frameToKnot[fkey(kn.as, kn.ak)] = e; topo.frameToKnot[fkey(kn.bs, kn.bk)] = e;
  }
}

// topology.js:177-201: over_s(k) = (k + φ_s) % 2 == 0, φ from a BFS over the knots
void assignParity(Topology& topo) {
  int n = topo.nStrands();
  std::vector<int> phi(n, -1); std::vector<std::vector<std::pair<int, int>>> adj(n);
  auto mod2 = [](int x) { return ((x % 2) + 2) % 2; };
  for (int e : topo.knotOrder) {
    const Knot& kn = topo.knots[e];
    int req = mod2(kn.ak + kn.bk + 1);
    adj[kn.as].push_back({kn.bs, req}); adj[kn.bs].push_back({kn.as, req});
  }
  for (int root = 0; root < n; root++) {
    if (phi[root] >= 0) continue;
    phi[root] = root == 0 ? mod2(topo.seedPhase) : 0;
    std::vector<int> dq{root}; size_t head = 0;
    while (head < dq.size()) {
      int s = dq[head++];
      for (auto [t, req] : adj[s]) { int want = mod2(req - phi[s]); if (phi[t] < 0) { phi[t] = want; dq.push_back(t); } }
    }
  }
  int conflicts = 0;
  for (int e : topo.knotOrder) {
    Knot& kn = topo.knots[e];
    bool overS = mod2(kn.ak + phi[kn.as]) == 0, overT = mod2(kn.bk + phi[kn.bs]) == 0;
    if (overS == overT) { kn.conflict = true; conflicts++; setOver(kn, 0); }
    else setOver(kn, overS ? 0 : 1);
  }
  topo.parityConflicts = conflicts;
}

void applyFlips(Topology& topo, const int32_t* flips, int nflips) {  // topology.js:203-211
  for (int i = 0; i < nflips; i++) {
    auto it = topo.frameToKnot.find(fkey(flips[2 * i], flips[2 * i + 1]));
    if (it == topo.frameToKnot.end()) continue;
    Knot& kn = topo.knots[it->second];
    setOver(kn, kn.overSide == 0 ? 1 : 0);
    kn.flipped = !kn.flipped;
  }
}

// topology.js:269-280: centripetal Catmull-Rom point between K[i1] and K[i2]
V3 catmullRom(const std::vector<V3>& K, int i0, int i1, int i2, int i3, double t) {
  const V3 P1 = K[i1], P2 = K[i2];
  const V3 P0 = i0 >= 0 ? K[i0] : sub3(scl3(P1, 2), P2);
  const V3 P3 = i3 >= 0 ? K[i3] : sub3(scl3(P2, 2), P1);
  auto tj = [](double ti, const V3& a, const V3& b) { return ti + std::sqrt(std::max(norm3(sub3(b, a)), 1e-12)); };
  double t0 = 0.0, t1 = tj(t0, P0, P1), t2 = tj(t1, P1, P2), t3 = tj(t2, P2, P3);
  double tt = t1 + t * (t2 - t1);
  auto lerp = [&](const V3& a, const V3& b, double ta, double tb) { return tb - ta < 1e-12 ? a : scl3(add3(scl3(a, tb - tt), scl3(b, tt - ta)), 1 / (tb - ta)); };
  V3 A1 = lerp(P0, P1, t0, t1), A2 = lerp(P1, P2, t1, t2), A3 = lerp(P2, P3, t2, t3);
  V3 B1 = lerp(A1, A2, t0, t2), B2 = lerp(A2, A3, t1, t3);
  return lerp(B1, B2, t1, t2);
}

// topology.js:283-292
std::vector<V3> nodeTangents(const std::vector<double>& P, bool closed) {
  int n = int(P.size() / 3);
  auto pt = [&](int i) { return V3{P[3 * i], P[3 * i + 1], P[3 * i + 2]}; };
  if (n == 1) return {V3{0, 0, 1}};
  std::vector<V3> T;
  if (closed) { for (int i = 0; i < n; i++) T.push_back(normalize3(sub3(pt((i + 1) % n), pt((i - 1 + n) % n)))); return T; }
  T.push_back(sub3(pt(1), pt(0)));
  for (int i = 1; i < n - 1; i++) T.push_back(add3(normalize3(sub3(pt(i + 1), pt(i))), normalize3(sub3(pt(i), pt(i - 1)))));
  T.push_back(sub3(pt(n - 1), pt(n - 2)));
  for (auto& t : T) t = normalize3(t);
  return T;
}

std::vector<double> nodeFramesOf(const std::vector<double>& P, const std::vector<V3>& hints, bool closed) {  // topology.js:294-300
  int n = int(P.size() / 3);
  std::vector<double> out(9 * n);
  if (n == 1) { double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}; std::copy(I, I + 9, out.begin()); return out; }
  auto T = nodeTangents(P, closed);
  for (int i = 0; i < n; i++) { auto R = frameFromTangentNormal(T[i], hints[i]); std::copy(R.begin(), R.end(), out.begin() + 9 * i); }
  return out;
}

// topology.js:214-266
void buildPolylines(Topology& topo) {
  const Mesh& mesh = *topo.mesh;
  auto mid = [&](int e) { return V3{mesh.edgeMid[3 * e], mesh.edgeMid[3 * e + 1], mesh.edgeMid[3 * e + 2]}; };
  auto en = [&](int e) { return V3{mesh.edgeN[3 * e], mesh.edgeN[3 * e + 1], mesh.edgeN[3 * e + 2]}; };
  double h = topo.halfHeight ? topo.thickness : 0.0;
  int sub = std::max(1, topo.subdivisions);
  for (int s = 0; s < topo.nStrands(); s++) {
    const Strand& st = topo.strands[s];
    int nk = int(st.edges.size());
    std::vector<V3> Nn(nk), Kp(nk), K0(nk); std::vector<double> sign(nk), off(nk);
    for (int k = 0; k < nk; k++) {
      int e = st.edges[k];
      Nn[k] = en(e);
      sign[k] = isOverFrame(topo.knots.at(e), s, k) ? 1.0 : -1.0;
      // a hybrid weave lifts each knot by the mean height of its two laths' groups (topology.js:225)
      const Knot& kn = topo.knots.at(e);
      const double he = (!topo.strandH.empty() && topo.halfHeight) ? 0.5 * (topo.strandH[kn.as] + topo.strandH[kn.bs]) : h;
      Kp[k] = add3(mid(e), scl3(Nn[k], 0.5 * he * sign[k]));
      K0[k] = mid(e);
      off[k] = (0.5 * h) * sign[k];
    }
    std::vector<V3> nodes, normals; std::vector<int> knotRows;
    if (!st.closed && topo.overhang > 0 && nk > 1) {
      V3 t0 = normalize3(sub3(Kp[1], Kp[0]));
      double seg = norm3(sub3(Kp[1], Kp[0])) / sub;
      int m = std::max(1, int(roundHalfEven(topo.overhang / std::max(seg, 1e-9))));
      for (int i = 0; i < m; i++) { double d = topo.overhang * (1 - double(i) / m); nodes.push_back(sub3(Kp[0], scl3(t0, d))); normals.push_back(Nn[0]); }
    }
    double step = 1.0 / sub;
    for (int k = 0; k < nk; k++) {
      knotRows.push_back(int(nodes.size()));
      nodes.push_back(Kp[k]); normals.push_back(Nn[k]);
      int nxt = k + 1 < nk ? k + 1 : (st.closed ? 0 : -1);
      if (nxt >= 0) {
        int prv = (st.closed || k > 0) ? (((k - 1) % nk) + nk) % nk : -1;
        int nx2 = (st.closed || nxt + 1 < nk) ? (nxt + 1) % nk : -1;
        for (int i = 1; i < sub; i++) {
          double t = i * step;
          V3 base = catmullRom(K0, prv, k, nxt, nx2, t);
          V3 nrm = normalize3(add3(scl3(Nn[k], 1 - t), scl3(Nn[nxt], t)));
          double w = 0.5 * (1.0 + std::cos(M_PI * t));
          double o = w * off[k] + (1 - w) * off[nxt];
          nodes.push_back(add3(base, scl3(nrm, o))); normals.push_back(nrm);
        }
      }
    }
    if (!st.closed && topo.overhang > 0 && nk > 1) {
      V3 t1 = normalize3(sub3(Kp[nk - 1], Kp[nk - 2]));
      double seg = norm3(sub3(Kp[nk - 1], Kp[nk - 2])) / sub;
      int m = std::max(1, int(roundHalfEven(topo.overhang / std::max(seg, 1e-9))));
      for (int i = 1; i <= m; i++) { nodes.push_back(add3(Kp[nk - 1], scl3(t1, topo.overhang * i / m))); normals.push_back(Nn[nk - 1]); }
    }
    std::vector<double> P(3 * nodes.size());
    for (size_t i = 0; i < nodes.size(); i++) for (int c = 0; c < 3; c++) P[3 * i + c] = nodes[i][c];
    topo.nodeFrames.push_back(nodeFramesOf(P, normals, st.closed));
    topo.polylines.push_back(std::move(P));
    for (int k = 0; k < nk; k++) topo.knotNode[fkey(s, k)] = knotRows[k];
  }
}

double interp1(double x, const std::vector<int>& xp, const std::vector<double>& fp) {  // topology.js:302-310
  if (x <= xp[0]) return fp[0];
  int last = int(xp.size()) - 1;
  if (x >= xp[last]) return fp[last];
  int lo = 0, hi = last;
  while (hi - lo > 1) { int m = (lo + hi) >> 1; if (xp[m] <= x) lo = m; else hi = m; }
  double slope = (fp[lo + 1] - fp[lo]) / double(xp[lo + 1] - xp[lo]);
  return slope * (x - xp[lo]) + fp[lo];
}

// topology.js:314-345: knot normals from the crossing of the centrelines, interpolated, frames rebuilt
void consistentFrames(Topology& topo) {
  if (!topo.mesh || topo.knots.empty()) return;
  const Mesh& mesh = *topo.mesh;
  int S = topo.nStrands();
  std::vector<std::vector<V3>> tang(S); std::vector<std::vector<std::pair<bool, V3>>> hints(S);
  for (int s = 0; s < S; s++) { tang[s] = nodeTangents(topo.polylines[s], topo.strands[s].closed); hints[s].assign(topo.polylines[s].size() / 3, {false, V3{0, 0, 0}}); }
  for (int e : topo.knotOrder) {
    const Knot& kn = topo.knots[e];
    int rA = topo.knotNode[fkey(kn.as, kn.ak)], rB = topo.knotNode[fkey(kn.bs, kn.bk)];
    V3 n = cross3(tang[kn.as][rA], tang[kn.bs][rB]);
    double nn = norm3(n); V3 ref{mesh.edgeN[3 * e], mesh.edgeN[3 * e + 1], mesh.edgeN[3 * e + 2]};
    n = nn < 1e-9 ? ref : scl3(n, 1 / nn);
    if (dot3(n, ref) < 0) n = scl3(n, -1);
    hints[kn.as][rA] = {true, n}; hints[kn.bs][rB] = {true, n};
  }
  for (int s = 0; s < S; s++) {
    const auto& H = hints[s]; std::vector<int> known;
    for (int i = 0; i < int(H.size()); i++) if (H[i].first) known.push_back(i);
    if (known.empty()) continue;
    std::vector<V3> Hf(H.size());
    if (known.size() == 1) { for (auto& v : Hf) v = H[known[0]].second; }
    else {
      std::vector<double> comp[3];
      for (int c = 0; c < 3; c++) for (int i : known) comp[c].push_back(H[i].second[c]);
      for (int i = 0; i < int(H.size()); i++) {
        if (i < known.front()) Hf[i] = H[known.front()].second;
        else if (i > known.back()) Hf[i] = H[known.back()].second;
        else for (int c = 0; c < 3; c++) Hf[i][c] = interp1(i, known, comp[c]);
      }
    }
    for (auto& v : Hf) { double l = std::max(norm3(v), 1e-12); v = V3{v[0] / l, v[1] / l, v[2] / l}; }
    topo.nodeFrames[s] = nodeFramesOf(topo.polylines[s], Hf, topo.strands[s].closed);
  }
}

// ------------------------------------------------------------------------------------------------ the model
struct Model {
  // topology kept for selectors
  int S = 0;
  std::vector<int> strandKnots;          // knots per strand (frame count)
  std::vector<uint8_t> strandClosedV;
  std::map<int64_t, int> nodeFrameIds;   // fkey → node
  std::vector<std::pair<int64_t, int64_t>> boundaryFrames;
  wsm_info info{};
  ws_params par{};
  int nPointLocal = 0;
  double lRef = 1.0;
  int n = 0, m = 0, p = 0, K = 0, C = 0;
  std::vector<int32_t> nodeOffset, segOffset, jointOffset, segI0, segI1, jSeg0, jSeg1, jNode, cSeg, cNode;
  std::vector<int32_t> kStrand, kSegOff, kNseg, plNode;
  std::vector<int32_t> kKnot;            // per crossing: mesh edge id, strand_a, k_a, strand_b, k_b, over strand [6K]
  std::vector<uint8_t> strandClosed, nodeFree, kEnabled, kLocked;
  std::vector<double> l0, jLen, EA, GA, EIx, EIy, GJ, omega0, cLen, cB, cQ, mass, krot, loadConst, plF;
  std::vector<double> kwk, khk, kfc90, kE90, kHs, kw, kh, nodeW, nodeH;
  // state
  std::vector<double> X, X0, V, Q, Q0, W, u, nPrev, slip, lam, penAcc, theta0, rotSlip, dent, climb, tanpsi, dentE, floorA;
  std::vector<uint8_t> stuck, floorState;
  std::vector<double> floorF;
  std::vector<double> N, Vsh, M, cM, reactions, Fn, Ft, gap, kstate, theta, Anom, pressure, util, Mrot, phi, mid;
  std::vector<int32_t> ctype, faceSide;
  int32_t iteration = 0; double kePrev = 0.0, residual = INFINITY;
};

struct Setup {
  wsm_scene sc;
  std::vector<wsm_anchor> anchors;
  std::vector<wsm_point_load> loads;
  wsm_prep prep{};
  bool hasPrep = false;
  double meshArea = 0;     // area of the welded mesh (a UDL per area spreads over it); 0 without a mesh
  std::vector<wsm_group> extra;        // lath groups 2, 3, … of a hybrid weave; group 1 is sc.grp (scene.js:53-56)
  std::vector<int32_t> strandGroup;    // the group of each strand; empty: every strand in group 1
  // pieces of a discretisation (Scene.joints, Scene.hybrid.pieces): each strand's joint arc lengths, ascending, and
  // the group assigned to piece p of strand s (key fkey(s, p)); pieceGroups: the scene assigns pieces at all
  std::vector<std::vector<double>> joints;
  std::map<int64_t, int> pieces;
  bool pieceGroups = false;
  bool valid(int g) const { return g >= 2 && g - 2 < int(extra.size()); }
  // scene.js:57-61: an assignment outside 2 … 1 + #extra falls back to group 1
  int groupIdOfStrand(int s) const {
    const int g = s < int(strandGroup.size()) ? strandGroup[s] : 1;
    return valid(g) ? g : 1;
  }
  // pieceGroup (scene.js:66-70): the piece's own assignment, else the strand's group
  int groupIdOfPiece(int s, int p) const {
    const auto it = pieces.find(fkey(s, p));
    const int g = it == pieces.end() ? 0 : it->second;
    return valid(g) ? g : groupIdOfStrand(s);
  }
  const wsm_group& group(int g) const { return valid(g) ? extra[g - 2] : sc.grp; }
  const wsm_group& groupOfStrand(int s) const { return group(groupIdOfStrand(s)); }
};

// np.searchsorted(a, v, 'right') on an ascending array (searchsorted, fields.js:88-93)
int searchsortedRight(const std::vector<double>& a, double v) {
  return int(std::upper_bound(a.begin(), a.end(), v) - a.begin());
}

// Σ of the triangle areas (Mesh.area, mesh.js:50-58)
double meshArea(const Mesh& mesh) {
  double a = 0;
  for (int f = 0; f < mesh.nF; f++) {
    const double* Vi = &mesh.V[3 * size_t(mesh.F[3 * f])];
    const double* Vj = &mesh.V[3 * size_t(mesh.F[3 * f + 1])];
    const double* Vk = &mesh.V[3 * size_t(mesh.F[3 * f + 2])];
    const double ux = Vj[0] - Vi[0], uy = Vj[1] - Vi[1], uz = Vj[2] - Vi[2];
    const double vx = Vk[0] - Vi[0], vy = Vk[1] - Vi[1], vz = Vk[2] - Vi[2];
    a += 0.5 * std::hypot(uy * vz - uz * vy, uz * vx - ux * vz, ux * vy - uy * vx);
  }
  return a;
}

// section stiffness (params.js:33-45) × beam factors (:49-57)
struct Stiff { double EA, EIx, EIy, GJ, GA; double act[3]; };
Stiff groupStiffness(const wsm_group& g) {
  auto rectJ = [](double a, double b) { if (a < b) std::swap(a, b); double r = b / a; return a * b * b * b * (1.0 / 3.0 - 0.21 * r * (1.0 - r * r * r * r / 12.0)); };
  double A = g.width * g.height;
  double Ix = g.Ix > 0 ? g.Ix : g.width * g.height * g.height * g.height / 12.0;
  double Iy = g.Iy > 0 ? g.Iy : g.height * g.width * g.width * g.width / 12.0;
  double J = g.J > 0 ? g.J : rectJ(g.width, g.height);
  Stiff k{g.E * A, g.E * Ix, g.E * Iy, g.G * J, 5.0 / 6.0 * g.G * A, {1, 1, 1}};
  auto c01 = [](double v) { return std::min(std::max(v, 0.0), 1.0); };
  if (g.beam_advanced) {
    k.EA *= g.k_axial; k.EIx *= g.k_flat; k.EIy *= g.k_edge; k.GJ *= g.k_twist; k.GA *= g.k_shear;
    k.act[0] = c01(g.act_flat); k.act[1] = c01(g.act_edge); k.act[2] = c01(g.act_twist);
  } else {
    k.EIx *= g.stiffness; k.EIy *= g.stiffness; k.GJ *= g.stiffness;
    k.act[0] = k.act[1] = k.act[2] = c01(g.activation);
  }
  return k;
}

// node tangents and mean node frames (PolylineContact._prepare, solver_cpu.js:179-193; nodeQuats, model.js:89-105)
void prepareNodes(const Model& md, std::vector<double>& Tn, std::vector<double>& Qn) {
  int n = md.n, m = md.m;
  const auto& X = md.X; const auto& Q = md.Q;
  Tn.assign(3 * n, 0.0);
  std::vector<double> eh(3 * m);
  for (int j = 0; j < m; j++) {
    int a = md.segI0[j], b = md.segI1[j];
    double ex = X[3 * b] - X[3 * a], ey = X[3 * b + 1] - X[3 * a + 1], ez = X[3 * b + 2] - X[3 * a + 2];
    double l = std::max(std::sqrt(ex * ex + ey * ey + ez * ez), 1e-12);
    eh[3 * j] = ex / l; eh[3 * j + 1] = ey / l; eh[3 * j + 2] = ez / l;
  }
  for (int j = 0; j < m; j++) { int a = md.segI0[j]; for (int k = 0; k < 3; k++) Tn[3 * a + k] += eh[3 * j + k]; }
  for (int j = 0; j < m; j++) { int b = md.segI1[j]; for (int k = 0; k < 3; k++) Tn[3 * b + k] += eh[3 * j + k]; }
  for (int i = 0; i < n; i++) {
    double l = std::max(std::sqrt(Tn[3 * i] * Tn[3 * i] + Tn[3 * i + 1] * Tn[3 * i + 1] + Tn[3 * i + 2] * Tn[3 * i + 2]), 1e-12);
    for (int k = 0; k < 3; k++) Tn[3 * i + k] /= l;
  }
  std::vector<int> ref(n, -1);
  for (int j = 0; j < m; j++) ref[md.segI1[j]] = j;
  for (int j = 0; j < m; j++) ref[md.segI0[j]] = j;
  Qn.assign(4 * n, 0.0);
  for (const auto* idx : {&md.segI0, &md.segI1}) {
    for (int j = 0; j < m; j++) {
      int i = (*idx)[j], r = ref[i];
      double d = Q[4 * j] * Q[4 * r] + Q[4 * j + 1] * Q[4 * r + 1] + Q[4 * j + 2] * Q[4 * r + 2] + Q[4 * j + 3] * Q[4 * r + 3];
      double sg = d < 0 ? -1 : 1;
      for (int c = 0; c < 4; c++) Qn[4 * i + c] += sg * Q[4 * j + c];
    }
  }
  for (int i = 0; i < n; i++) qnormalize(&Qn[4 * i]);
}

struct Eval { std::vector<double> P, t, L; };
// PolylineContact._segment (solver_cpu.js:144-149)
inline void segmentOf(const Model& md, int side, int c, int& seg, double& t) {
  double u = md.u[2 * c + side]; int n = md.kNseg[2 * c + side];
  int i = int(std::min(std::max(std::floor(u), 0.0), double(std::max(n - 1, 0))));
  t = std::min(std::max(u - i, 0.0), 1.0);
  seg = md.kSegOff[2 * c + side] + i;
}
// PolylineContact._eval (solver_cpu.js:197-228): C¹ Hermite point and tangent (frames are not needed at setup)
void evalSide(const Model& md, const std::vector<double>& Tn, int side, Eval& ev) {
  int K = md.K; ev.P.assign(3 * K, 0); ev.t.assign(3 * K, 0); ev.L.assign(K, 0);
  const auto& X = md.X;
  for (int c = 0; c < K; c++) {
    int seg; double t; segmentOf(md, side, c, seg, t);
    int a = md.segI0[seg], b = md.segI1[seg];
    double dx = X[3 * b] - X[3 * a], dy = X[3 * b + 1] - X[3 * a + 1], dz = X[3 * b + 2] - X[3 * a + 2];
    double L = std::sqrt(dx * dx + dy * dy + dz * dz);
    double t2 = t * t, t3 = t * t * t;
    double h00 = 2 * t3 - 3 * t2 + 1, h10 = t3 - 2 * t2 + t, h01 = -2 * t3 + 3 * t2, h11 = t3 - t2;
    double d00 = 6 * t2 - 6 * t, d10 = 3 * t2 - 4 * t + 1, d01 = -6 * t2 + 6 * t, d11 = 3 * t2 - 2 * t;
    double tg[3];
    for (int k = 0; k < 3; k++) {
      ev.P[3 * c + k] = h00 * X[3 * a + k] + (h10 * L) * Tn[3 * a + k] + h01 * X[3 * b + k] + (h11 * L) * Tn[3 * b + k];
      tg[k] = d00 * X[3 * a + k] + (d10 * L) * Tn[3 * a + k] + d01 * X[3 * b + k] + (d11 * L) * Tn[3 * b + k];
    }
    double tl = std::max(std::sqrt(tg[0] * tg[0] + tg[1] * tg[1] + tg[2] * tg[2]), 1e-12);
    for (int k = 0; k < 3; k++) ev.t[3 * c + k] = tg[k] / tl;
    ev.L[c] = L;
  }
}

// PolylineContact.initialize (solver_cpu.js:99-139): stuck points paired at the closest points of the centrelines
void contactInitialize(Model& md, int sweeps = 3) {
  int K = md.K;
  std::vector<double> Tn, Qn; Eval A, B;
  auto offsets = [&](std::vector<double>& St, std::vector<double>& nOut) {
    prepareNodes(md, Tn, Qn); evalSide(md, Tn, 0, A); evalSide(md, Tn, 1, B);
    St.assign(3 * K, 0); nOut.assign(3 * K, 0);
    for (int c = 0; c < K; c++) {
      const double* tA = &A.t[3 * c]; const double* tB = &B.t[3 * c];
      double nx = tA[1] * tB[2] - tA[2] * tB[1], ny = tA[2] * tB[0] - tA[0] * tB[2], nz = tA[0] * tB[1] - tA[1] * tB[0];
      double nn = std::sqrt(nx * nx + ny * ny + nz * nz);
      if (nn < 1e-6) { nx = md.nPrev[3 * c]; ny = md.nPrev[3 * c + 1]; nz = md.nPrev[3 * c + 2]; }
      else { double d = std::max(nn, 1e-12); nx /= d; ny /= d; nz /= d; }
      if (nx * md.nPrev[3 * c] + ny * md.nPrev[3 * c + 1] + nz * md.nPrev[3 * c + 2] < 0) { nx = -nx; ny = -ny; nz = -nz; }
      double S0 = B.P[3 * c] - A.P[3 * c], S1 = B.P[3 * c + 1] - A.P[3 * c + 1], S2 = B.P[3 * c + 2] - A.P[3 * c + 2];
      double g = S0 * nx + S1 * ny + S2 * nz;
      St[3 * c] = S0 - g * nx; St[3 * c + 1] = S1 - g * ny; St[3 * c + 2] = S2 - g * nz;
      nOut[3 * c] = nx; nOut[3 * c + 1] = ny; nOut[3 * c + 2] = nz;
    }
  };
  auto nrm = [](const std::vector<double>& St, int c) { return std::sqrt(St[3 * c] * St[3 * c] + St[3 * c + 1] * St[3 * c + 1] + St[3 * c + 2] * St[3 * c + 2]); };
  std::vector<double> St, nn, St2;
  for (int it = 0; it < sweeps; it++) {
    offsets(St, nn);
    std::vector<double> uOld = md.u, before(K);
    for (int c = 0; c < K; c++) {
      const double* tA = &A.t[3 * c]; const double* tB = &B.t[3 * c];
      double taa = tA[0] * tA[0] + tA[1] * tA[1] + tA[2] * tA[2], tbb = tB[0] * tB[0] + tB[1] * tB[1] + tB[2] * tB[2], tab = tA[0] * tB[0] + tA[1] * tB[1] + tA[2] * tB[2];
      double sa = St[3 * c] * tA[0] + St[3 * c + 1] * tA[1] + St[3 * c + 2] * tA[2], sb = St[3 * c] * tB[0] + St[3 * c + 1] * tB[1] + St[3 * c + 2] * tB[2];
      double det = taa * tbb - tab * tab; bool ok = det > 1e-3 && md.kEnabled[c];
      double ca = ok ? (sa * tbb - sb * tab) / det : 0, cb = ok ? (sb * taa - sa * tab) / det : 0;
      before[c] = nrm(St, c);
      md.u[2 * c] = std::min(std::max(md.u[2 * c] + ca / std::max(A.L[c], 1e-12), 0.0), double(md.kNseg[2 * c]));
      md.u[2 * c + 1] = std::min(std::max(md.u[2 * c + 1] - cb / std::max(B.L[c], 1e-12), 0.0), double(md.kNseg[2 * c + 1]));
    }
    offsets(St2, nn);
    for (int c = 0; c < K; c++) if (nrm(St2, c) >= before[c]) { md.u[2 * c] = uOld[2 * c]; md.u[2 * c + 1] = uOld[2 * c + 1]; }
  }
  offsets(St, nn);
  md.nPrev = nn;
}

// ------------------------------------------------------------------------------------------------ CpuSolver setup
struct Built {
  int S = 0;
  std::vector<std::vector<double>> polylines, nodeFrames;
  std::vector<uint8_t> closed;
  std::vector<int> nKnots;                                // frames per strand
  struct KnotRec { int as, ak, bs, bk; bool isBoundary; V3 edgeNormal; int edge = 0, overSide = 0; };
  std::vector<KnotRec> knots;                             // sorted by knot id (edge id)
  std::map<int64_t, int> knotNode;                        // fkey → row
  double overhang = 0;
  // the mesh graph, for the 'region' selector: vertices, edges, and the knots' edges in the topology's insertion order
  std::vector<double> meshV;
  std::vector<int32_t> meshE;
  std::vector<int> knotOrder;                             // indices into `knots`
};

// regionFrames (model.js:127-145): the mesh graph distance from the two vertices of the seed frame's knot edge
// (Dijkstra over the edges, lengths in scene units), then both frames of every knot whose edge's two vertex
// distances average at most r, in the topology's knot order
std::vector<int64_t> regionFrames(const Built& tp, int s0, int k0, double r) {
  std::vector<int64_t> out;
  const int nV = int(tp.meshV.size() / 3), nE = int(tp.meshE.size() / 2);
  if (nE == 0) return out;
  int seed = -1;
  for (const auto& kr : tp.knots)
    if ((kr.as == s0 && kr.ak == k0) || (kr.bs == s0 && kr.bk == k0)) { seed = kr.edge; break; }
  if (seed < 0 || seed >= nE) return out;
  std::vector<std::vector<std::pair<int, double>>> nbr(nV);
  for (int e = 0; e < nE; e++) {
    const int a = tp.meshE[2 * e], b = tp.meshE[2 * e + 1];
    const double l = std::hypot(tp.meshV[3 * a] - tp.meshV[3 * b], tp.meshV[3 * a + 1] - tp.meshV[3 * b + 1],
                                tp.meshV[3 * a + 2] - tp.meshV[3 * b + 2]);
    nbr[a].push_back({b, l}); nbr[b].push_back({a, l});
  }
  std::vector<double> dist(nV, INFINITY);
  using Item = std::pair<double, int>;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
  for (int v : {tp.meshE[2 * seed], tp.meshE[2 * seed + 1]}) { dist[v] = 0.0; heap.push({0.0, v}); }
  while (!heap.empty()) {
    const auto [d, v] = heap.top(); heap.pop();
    if (d > dist[v]) continue;
    for (const auto& [w, l] : nbr[v]) if (d + l < dist[w]) { dist[w] = d + l; heap.push({d + l, w}); }
  }
  for (int q : tp.knotOrder) {
    const auto& kr = tp.knots[q];
    if (0.5 * (dist[tp.meshE[2 * kr.edge]] + dist[tp.meshE[2 * kr.edge + 1]]) <= r) {
      out.push_back(fkey(kr.as, kr.ak)); out.push_back(fkey(kr.bs, kr.bk));
    }
  }
  return out;
}

std::unique_ptr<Model> buildModel(const Built& tp, const Setup& su) {
  auto md = std::make_unique<Model>();
  const wsm_scene& sc = su.sc;
  int S = tp.S; md->S = S;
  // ---- buildState (model.js:10-75)
  md->nodeOffset.assign(S + 1, 0);
  for (int s = 0; s < S; s++) md->nodeOffset[s + 1] = md->nodeOffset[s] + int(tp.polylines[s].size() / 3);
  int n = md->nodeOffset[S]; md->n = n;
  md->X.assign(3 * n, 0);
  for (int s = 0; s < S; s++) std::copy(tp.polylines[s].begin(), tp.polylines[s].end(), md->X.begin() + 3 * md->nodeOffset[s]);
  std::vector<int> segStrand, segCounts;
  std::vector<double> Qv;
  md->segOffset.assign(S + 1, 0); md->jointOffset.assign(S + 1, 0);
  for (int s = 0; s < S; s++) {
    int off = md->nodeOffset[s], nn = int(tp.polylines[s].size() / 3);
    bool closed = tp.closed[s];
    int nseg = closed ? nn : nn - 1;
    int base = md->segOffset[s];
    const auto& P = tp.polylines[s]; const auto& R = tp.nodeFrames[s];
    for (int k = 0; k < nseg; k++) {
      int a = k, b = (k + 1) % nn;
      md->segI0.push_back(off + a); md->segI1.push_back(off + b); segStrand.push_back(s);
      V3 d{P[3 * b] - P[3 * a], P[3 * b + 1] - P[3 * a + 1], P[3 * b + 2] - P[3 * a + 2]};
      md->l0.push_back(std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
      V3 hint{R[9 * a + 3] + R[9 * b + 3], R[9 * a + 4] + R[9 * b + 4], R[9 * a + 5] + R[9 * b + 5]};
      auto Fr = frameFromTangentNormal(d, hint); double q[4]; fromFrame(Fr.data(), q);
      for (double v : q) Qv.push_back(v);
    }
    if (closed) for (int k = 0; k < nseg; k++) { md->jSeg0.push_back(base + k); md->jSeg1.push_back(base + (k + 1) % nseg); md->jNode.push_back(off + (k + 1) % nn); }
    else for (int k = 0; k < std::max(nseg - 1, 0); k++) { md->jSeg0.push_back(base + k); md->jSeg1.push_back(base + k + 1); md->jNode.push_back(off + 1 + k); }
    md->segOffset[s + 1] = base + nseg;
    md->jointOffset[s + 1] = int(md->jSeg0.size());
  }
  int m = md->segOffset[S], p = int(md->jSeg0.size());
  md->m = m; md->p = p;
  md->Q = Qv; md->Q0 = Qv;
  md->jLen.assign(p, 0);
  for (int j = 0; j < p; j++) md->jLen[j] = 0.5 * (md->l0[md->jSeg0[j]] + md->l0[md->jSeg1[j]]);
  md->strandClosed = tp.closed;
  md->strandKnots = tp.nKnots;
  for (auto& kv : tp.knotNode) { int s = int(kv.first / 1048576); md->nodeFrameIds[kv.first] = md->nodeOffset[s] + kv.second; }
  int K = int(tp.knots.size()); md->K = K;
  md->u.assign(2 * K, 0); md->nPrev.assign(3 * K, 0);
  for (int c = 0; c < K; c++) {
    const auto& kn = tp.knots[c];
    int fs[2][2] = {{kn.as, kn.ak}, {kn.bs, kn.bk}};
    for (int side = 0; side < 2; side++) {
      int s = fs[side][0]; int row = tp.knotNode.at(fkey(fs[side][0], fs[side][1])); int nseg = md->segOffset[s + 1] - md->segOffset[s];
      md->u[2 * c + side] = std::min(row, nseg);
    }
    int ia = md->nodeFrameIds.at(fkey(kn.as, kn.ak)), ib = md->nodeFrameIds.at(fkey(kn.bs, kn.bk));
    V3 d{md->X[3 * ib] - md->X[3 * ia], md->X[3 * ib + 1] - md->X[3 * ia + 1], md->X[3 * ib + 2] - md->X[3 * ia + 2]};
    double nd = norm3(d);
    for (int k = 0; k < 3; k++) md->nPrev[3 * c + k] = nd > 1e-12 ? d[k] / nd : kn.edgeNormal[k];
    if (kn.isBoundary) md->boundaryFrames.push_back({fkey(kn.as, kn.ak), fkey(kn.bs, kn.bk)});
  }
  md->V.assign(3 * n, 0); md->W.assign(3 * m, 0); md->omega0.assign(3 * p, 0);
  md->slip.assign(K, 0); md->stuck.assign(K, 1); md->lam.assign(K, 0);

  // ---- synthetic preparation: seeded out-of-plane noise on the initial nodes (rng.js generator)
  if (su.hasPrep && su.prep.noise > 0) {
    Rng rng(su.prep.seed);
    if (su.prep.wavelength > 0) {
      const int NW = 8;
      double th[NW], lam[NW], ph[NW], w[NW], ws = 0;
      for (int k = 0; k < NW; k++) {
        th[k] = 2 * M_PI * rng.uniform(); lam[k] = su.prep.wavelength * (1.0 + 3.0 * rng.uniform());
        ph[k] = 2 * M_PI * rng.uniform(); w[k] = 0.5 + rng.uniform(); ws += w[k];
      }
      for (int i = 0; i < n; i++) {
        double x = md->X[3 * i], y = md->X[3 * i + 1], z = 0;
        for (int k = 0; k < NW; k++) z += w[k] * std::sin(2 * M_PI * (x * std::cos(th[k]) + y * std::sin(th[k])) / lam[k] + ph[k]);
        md->X[3 * i + 2] += su.prep.noise * z / ws;
      }
    } else {
      for (int i = 0; i < n; i++) md->X[3 * i + 2] += (2.0 * rng.uniform() - 1.0) * su.prep.noise;
    }
  }
  md->X0 = md->X;

  // the strand of every segment and node, whose lath group sets its section, stiffness and weight
  std::vector<int> segS(m), nodeS(n);
  for (int s = 0; s < S; s++) {
    for (int j = md->segOffset[s]; j < md->segOffset[s + 1]; j++) segS[j] = s;
    for (int i = md->nodeOffset[s]; i < md->nodeOffset[s + 1]; i++) nodeS[i] = s;
  }
  // the lath group of every segment and node (Scene.segmentGroups / nodeGroups, scene.js:75-90): the strand's, or,
  // where the scene assigns pieces of a discretisation, the group of the piece that holds the segment's midpoint /
  // the node by arc length along the strand on the built positions (segmentPieceIndex / pieceIndex,
  // discretize.js:334-358; a strand without joints is piece 0 throughout)
  std::vector<int> gSeg(m), gNode(n);
  if (!su.pieceGroups) {
    for (int j = 0; j < m; j++) gSeg[j] = su.groupIdOfStrand(segS[j]);
    for (int i = 0; i < n; i++) gNode[i] = su.groupIdOfStrand(nodeS[i]);
  } else {
    for (int s = 0; s < S; s++) {
      const int i0 = md->nodeOffset[s], nn = md->nodeOffset[s + 1] - i0;
      const std::vector<double> none;
      const std::vector<double>& js = s < int(su.joints.size()) ? su.joints[s] : none;
      std::vector<double> Sarc(nn, 0.0);
      const auto& X = md->X;
      for (int i = 1; i < nn; i++) {
        const int a = 3 * (i0 + i - 1), b = 3 * (i0 + i);
        Sarc[i] = Sarc[i - 1] + std::hypot(X[b] - X[a], X[b + 1] - X[a + 1], X[b + 2] - X[a + 2]);
      }
      for (int j = md->segOffset[s]; j < md->segOffset[s + 1]; j++) {
        int pc = 0;
        if (!js.empty()) {
          const int a = md->segI0[j], b = md->segI1[j];
          const double len = std::hypot(X[3 * b] - X[3 * a], X[3 * b + 1] - X[3 * a + 1], X[3 * b + 2] - X[3 * a + 2]);
          pc = searchsortedRight(js, Sarc[a - i0] + 0.5 * len);
        }
        gSeg[j] = su.groupIdOfPiece(s, pc);
      }
      for (int i = 0; i < nn; i++) gNode[i0 + i] = su.groupIdOfPiece(s, js.empty() ? 0 : searchsortedRight(js, Sarc[i]));
    }
  }
  // ---- strandStiffness (model.js:236-252): each segment takes its group, each joint the activation of the group of
  // its first segment
  md->EA.resize(m); md->GA.resize(m); md->EIx.resize(m); md->EIy.resize(m); md->GJ.resize(m);
  std::vector<Stiff> segK(m);
  for (int j = 0; j < m; j++) {
    const Stiff k = groupStiffness(su.group(gSeg[j]));
    segK[j] = k; md->EA[j] = k.EA; md->GA[j] = k.GA; md->EIx[j] = k.EIx; md->EIy[j] = k.EIy; md->GJ[j] = k.GJ;
  }
  // ---- CosseratRod (rod_cosserat.js:49-72): Ω0 = (1 − act)·asBuilt(Q0)
  for (int j = 0; j < p; j++) {
    const Stiff& k = segK[md->jSeg0[j]];
    double act[3] = {k.act[0], k.act[1], k.act[2]};
    if (sc.prestress_asbuilt) act[0] = act[1] = act[2] = 0.0;
    if (act[0] < 1.0 || act[1] < 1.0 || act[2] < 1.0) {
      double psi[3];
      relRotation(&md->Q0[4 * md->jSeg0[j]], &md->Q0[4 * md->jSeg1[j]], psi);
      for (int c = 0; c < 3; c++) md->omega0[3 * j + c] = (1.0 - act[c]) * (psi[c] / md->jLen[j]);
    }
  }
  // ---- Loads (model.js:191-231)
  md->loadConst.assign(3 * n, 0);
  std::vector<double> trib(n, 0);
  for (int j = 0; j < m; j++) trib[md->segI0[j]] += 0.5 * md->l0[j];
  for (int j = 0; j < m; j++) trib[md->segI1[j]] += 0.5 * md->l0[j];
  if (sc.gravity) {
    double g[3] = {sc.gravity_dir[0], sc.gravity_dir[1], sc.gravity_dir[2]};
    double gl = std::max(std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]), 1e-12);
    for (double& v : g) v /= gl;
    double L = sc.length_to_m, F = sc.force_to_n;
    for (int j = 0; j < m; j++) {
      // self-weight of the segment's group (model.js:203-211; units.js:21-25)
      const wsm_group& gr = su.group(gSeg[j]);
      const double areaSI = gr.width * gr.height * L * L, qSI = gr.density * areaSI * 9.80665, qline = qSI / F * L;
      double q = 0.5 * md->l0[j] * qline;
      for (int i : {md->segI0[j], md->segI1[j]}) for (int c = 0; c < 3; c++) md->loadConst[3 * i + c] += q * g[c];
    }
  }
  if (sc.udl == 2) {
    // per area: the mesh's area shared out in proportion to the nodes' tributary lengths (model.js:216-219)
    double total = 0;
    for (int i = 0; i < n; i++) total += trib[i];
    for (int i = 0; i < n; i++) {
      const double w = trib[i] / std::max(total, 1e-12) * su.meshArea;
      for (int c = 0; c < 3; c++) md->loadConst[3 * i + c] += w * sc.udl_q[c];
    }
  } else if (sc.udl) {
    for (int i = 0; i < n; i++) for (int c = 0; c < 3; c++) md->loadConst[3 * i + c] += trib[i] * sc.udl_q[c];
  }
  for (const auto& pl : su.loads) {
    auto it = md->nodeFrameIds.find(fkey(pl.s, pl.k));
    if (it == md->nodeFrameIds.end()) continue;
    if (pl.local) { md->plNode.push_back(it->second); for (int c = 0; c < 3; c++) md->plF.push_back(pl.F[c]); }
    else for (int c = 0; c < 3; c++) md->loadConst[3 * it->second + c] += pl.F[c];
  }
  md->nPointLocal = int(md->plNode.size());
  // ---- Anchors (model.js:150-189): union of fixities per frame in resolution order
  std::vector<std::pair<int64_t, std::array<bool, 6>>> masks;
  auto addMask = [&](int64_t key, const int32_t* fix) {
    for (auto& mk : masks) if (mk.first == key) { for (int i = 0; i < 6; i++) mk.second[i] = mk.second[i] || fix[i]; return; }
    std::array<bool, 6> a{}; for (int i = 0; i < 6; i++) a[i] = fix[i] != 0; masks.push_back({key, a});
  };
  for (const auto& an : su.anchors) {
    std::vector<int64_t> keys;
    switch (an.type) {
      case WSM_FRAME: if (md->nodeFrameIds.count(fkey(an.s, an.k))) keys.push_back(fkey(an.s, an.k)); break;
      case WSM_ALL_START: for (int s = 0; s < S; s++) if (!tp.closed[s]) keys.push_back(fkey(s, 0)); break;
      case WSM_ALL_END: for (int s = 0; s < S; s++) if (!tp.closed[s]) keys.push_back(fkey(s, tp.nKnots[s] - 1)); break;
      case WSM_ALL_ENDS:
        for (int s = 0; s < S; s++) if (!tp.closed[s]) keys.push_back(fkey(s, 0));
        for (int s = 0; s < S; s++) if (!tp.closed[s]) keys.push_back(fkey(s, tp.nKnots[s] - 1));
        break;
      case WSM_STRAND: if (an.s >= 0 && an.s < S) for (int kk = 0; kk < tp.nKnots[an.s]; kk++) keys.push_back(fkey(an.s, kk)); break;
      case WSM_BOUNDARY: for (auto& bf : md->boundaryFrames) { keys.push_back(bf.first); keys.push_back(bf.second); } break;
      case WSM_REGION: keys = regionFrames(tp, an.s, an.k, an.radius); break;
      default: throw std::runtime_error("unknown selector type");
    }
    for (int64_t key : keys) addMask(key, an.fix);
  }
  md->nodeFree.assign(3 * n, 1);
  std::vector<int> fixedNodes; std::vector<std::tuple<int, int, std::array<bool, 3>>> rotClamps;
  std::vector<std::vector<int>> adj(n);
  for (int j = 0; j < m; j++) { adj[md->segI0[j]].push_back(j); adj[md->segI1[j]].push_back(j); }
  for (auto& mk : masks) {
    auto it = md->nodeFrameIds.find(mk.first);
    if (it == md->nodeFrameIds.end()) continue;
    int nid = it->second; const auto& msk = mk.second;
    for (int c = 0; c < 3; c++) if (msk[c]) md->nodeFree[3 * nid + c] = 0;
    if ((msk[0] || msk[1] || msk[2]) && std::find(fixedNodes.begin(), fixedNodes.end(), nid) == fixedNodes.end()) fixedNodes.push_back(nid);
    for (int j : adj[nid]) if (msk[3] || msk[4] || msk[5]) rotClamps.push_back({nid, j, {msk[3], msk[4], msk[5]}});
  }
  // ---- setClamps (rod_cosserat.js:92-100)
  md->C = int(rotClamps.size());
  for (auto& [nid, j, mk] : rotClamps) {
    md->cNode.push_back(nid); md->cSeg.push_back(j); md->cLen.push_back(0.5 * md->l0[j]);
    md->cB.push_back(mk[0] ? md->EIx[j] : 0); md->cB.push_back(mk[1] ? md->EIy[j] : 0); md->cB.push_back(mk[2] ? md->GJ[j] : 0);
    for (int c = 0; c < 4; c++) md->cQ.push_back(md->Q[4 * j + c]);
  }
  // ---- PolylineContact (solver_cpu.js:62-96)
  ws_params par = sc.par;
  par.contacts = (sc.contacts && K > 0) ? 1 : 0;
  if (par.dent_after < 0) par.dent_after = par.warmup_iter + par.ramp_iter;   // solver_cpu.js:686
  md->kStrand.assign(2 * K, 0); md->kSegOff.assign(2 * K, 0); md->kNseg.assign(2 * K, 0);
  md->kwk.assign(2 * K, 0); md->khk.assign(2 * K, 0);
  md->kfc90.assign(2 * K, 0); md->kE90.assign(2 * K, 0); md->kHs.assign(2 * K, 0);
  md->kw.assign(K, 0); md->kh.assign(K, 0);
  md->kEnabled.assign(K, 1); md->kLocked.assign(K, sc.locked ? 1 : 0);
  md->kKnot.assign(6 * size_t(K), 0);
  for (int c = 0; c < K; c++) {
    const auto& kr = tp.knots[c];
    const int32_t row[6] = {kr.edge, kr.as, kr.ak, kr.bs, kr.bk, kr.overSide ? kr.bs : kr.as};
    std::copy(row, row + 6, md->kKnot.begin() + 6 * size_t(c));
  }
  for (int c = 0; c < K; c++) {
    md->kStrand[2 * c] = tp.knots[c].as; md->kStrand[2 * c + 1] = tp.knots[c].bs;
    for (int side = 0; side < 2; side++) {
      const int i = 2 * c + side, s = md->kStrand[i];
      md->kSegOff[i] = md->segOffset[s]; md->kNseg[i] = md->segOffset[s + 1] - md->segOffset[s];
      // each side's section and material: the group at the side's crossing node, its strand's or its piece's
      // (solver_cpu.js:66-78, 683)
      const auto kn = md->nodeFrameIds.find(side ? fkey(tp.knots[c].bs, tp.knots[c].bk) : fkey(tp.knots[c].as, tp.knots[c].ak));
      const wsm_group& gr = su.group(kn != md->nodeFrameIds.end() ? gNode[kn->second] : su.groupIdOfStrand(s));
      md->kwk[i] = gr.width; md->khk[i] = gr.height; md->kfc90[i] = gr.fc90; md->kE90[i] = gr.E90; md->kHs[i] = gr.H;
    }
    md->kw[c] = 0.5 * (md->kwk[2 * c] + md->kwk[2 * c + 1]); md->kh[c] = 0.5 * (md->khk[2 * c] + md->khk[2 * c + 1]);
    if (tp.overhang <= 0 && tp.knots[c].isBoundary) md->kEnabled[c] = 0;
    md->kw[c] = 0.5 * (md->kwk[2 * c] + md->kwk[2 * c + 1]); md->kh[c] = 0.5 * (md->khk[2 * c] + md->khk[2 * c + 1]);
  }
  md->theta0.assign(K, NAN); md->rotSlip.assign(K, 0); md->dent.assign(2 * K, 0); md->dentE.assign(K, 0); md->climb.assign(2 * K, 0);
  md->tanpsi.assign(2 * K, 1); md->penAcc.assign(K, 0);
  md->Fn.assign(K, 0); md->Ft.assign(K, 0); md->gap.assign(K, 0); md->kstate.assign(K, 0); md->theta.assign(K, 0); md->Anom.assign(K, 0);
  md->pressure.assign(K, 0); md->util.assign(K, 0); md->Mrot.assign(K, 0); md->phi.assign(2 * K, 0); md->mid.assign(3 * K, 0);
  md->ctype.assign(K, 0); md->faceSide.assign(K, -1);
  if (par.contacts) contactInitialize(*md);
  // ---- _buildGoals masses and inertias (solver_cpu.js:694-708)
  std::vector<double> kNode(n, 0.0), kSeg(m, 0.0);
  for (int j = 0; j < m; j++) { double kk = (md->EA[j] + md->GA[j]) / md->l0[j]; kNode[md->segI0[j]] += kk; kNode[md->segI1[j]] += kk; }  // rod_cosserat.js:186-191
  if (par.contacts) {
    double kc = par.strength + par.stick_strength;                                                    // solver_cpu.js:151-167
    for (int side = 0; side < 2; side++) {
      for (int c = 0; c < K; c++) { int seg; double t; segmentOf(*md, side, c, seg, t); kNode[md->segI0[seg]] += kc; }
      for (int c = 0; c < K; c++) { int seg; double t; segmentOf(*md, side, c, seg, t); kNode[md->segI1[seg]] += kc; }
    }
    if (par.rot_friction) {
      for (int side = 0; side < 2; side++) for (int pass = 0; pass < 2; pass++) for (int c = 0; c < K; c++) {
        int seg; double t; segmentOf(*md, side, c, seg, t);
        double kr = par.stick_strength * (0.38 * 0.38) * md->kwk[2 * c] * md->kwk[2 * c + 1] / 0.5;
        double l = std::max(md->l0[seg], 1e-12);
        kNode[pass ? md->segI1[seg] : md->segI0[seg]] += kr / (l * l);
      }
    }
    double tf = std::max(par.torsion_factor, 1e-3);                                                    // addTwistStiffness, ramp 1, Fn 0
    for (int side = 0; side < 2; side++) for (int c = 0; c < K; c++) {
      double kth = tf * ((0.0 + par.strength * 0.25 * md->kh[c]) * (0.5 * md->kw[c]) / 0.03 + par.stick_strength * (0.5 * md->kh[c]) * (0.5 * md->kh[c]));
      int seg; double t; segmentOf(*md, side, c, seg, t); kSeg[seg] += kth;
    }
  }
  if (par.floor_on) for (int i = 0; i < n; i++) kNode[i] += par.strength + par.stick_strength;        // Floor.addStiffness
  double ms = par.mass_scale;
  md->mass.assign(n, 0);
  for (int i = 0; i < n; i++) md->mass[i] = ms * std::max(kNode[i], 1e-12);
  md->krot.assign(3 * m, 0.0);                                                                          // addRotStiffness (rod_cosserat.js:196-209)
  for (int j = 0; j < p; j++) {
    double lb = std::max(md->jLen[j], 1e-12); int s0 = md->jSeg0[j];
    double Bv[3] = {md->EIx[s0], md->EIy[s0], md->GJ[s0]};
    for (int s : {md->jSeg0[j], md->jSeg1[j]}) for (int c = 0; c < 3; c++) md->krot[3 * s + c] += Bv[c] / lb;
  }
  for (int c = 0; c < md->C; c++) for (int k2 = 0; k2 < 3; k2++) md->krot[3 * md->cSeg[c] + k2] += md->cB[3 * c + k2] / std::max(md->cLen[c], 1e-12);
  for (int j = 0; j < m; j++) {
    int a = md->segI0[j], b = md->segI1[j];
    double L = std::hypot(md->X[3 * b] - md->X[3 * a], md->X[3 * b + 1] - md->X[3 * a + 1], md->X[3 * b + 2] - md->X[3 * a + 2]);
    double kk = md->GA[j] * L * L / md->l0[j];
    md->krot[3 * j] += kk; md->krot[3 * j + 1] += kk;
  }
  double lmin = INFINITY;
  for (double v : md->jLen) lmin = std::min(lmin, v);
  for (double v : md->l0) lmin = std::min(lmin, v);
  md->lRef = std::isfinite(lmin) ? lmin : 1.0;
  md->par = par;
  // floor goal sections and stick points (Floor constructor, solver_cpu.js:536-554)
  md->nodeW.resize(n); md->nodeH.resize(n);
  for (int i = 0; i < n; i++) { const wsm_group& gr = su.group(gNode[i]); md->nodeW[i] = gr.width; md->nodeH[i] = gr.height; }
  md->floorA.assign(2 * n, 0);
  for (int i = 0; i < n; i++) { md->floorA[2 * i] = md->X[3 * i]; md->floorA[2 * i + 1] = md->X[3 * i + 1]; }
  md->floorF.assign(3 * n, 0); md->floorState.assign(n, 0);
  // outputs
  md->N.assign(m, 0); md->Vsh.assign(2 * m, 0); md->M.assign(3 * p, 0); md->cM.assign(2 * md->C, 0); md->reactions.assign(3 * n, 0);
  md->info.n = n; md->info.m = m; md->info.p = p; md->info.K = K; md->info.C = md->C; md->info.S = S;
  double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
  for (int i = 0; i < n; i++) for (int c = 0; c < 2; c++) { lo[c] = std::min(lo[c], md->X[3 * i + c]); hi[c] = std::max(hi[c], md->X[3 * i + c]); }
  md->info.span_plan = n ? std::max(hi[0] - lo[0], hi[1] - lo[1]) : 0.0;
  return md;
}

Built fromTopology(const Topology& topo) {
  Built b; b.S = topo.nStrands(); b.polylines = topo.polylines; b.nodeFrames = topo.nodeFrames; b.overhang = topo.overhang;
  for (auto& s : topo.strands) { b.closed.push_back(s.closed ? 1 : 0); b.nKnots.push_back(int(s.edges.size())); }
  b.knotNode = topo.knotNode;
  for (auto& kv : topo.knots) {  // sorted edge order (model.js:50)
    const Knot& kn = kv.second;
    V3 en{topo.mesh->edgeN[3 * kn.edge], topo.mesh->edgeN[3 * kn.edge + 1], topo.mesh->edgeN[3 * kn.edge + 2]};
    b.knots.push_back({kn.as, kn.ak, kn.bs, kn.bk, kn.isBoundary, en, kn.edge, kn.overSide});
  }
  b.meshV = topo.mesh->V; b.meshE = topo.mesh->E;
  std::map<int, int> row;
  for (int q = 0; q < int(b.knots.size()); q++) row[b.knots[q].edge] = q;
  for (int e : topo.knotOrder) b.knotOrder.push_back(row.at(e));
  return b;
}

ws_model describe(Model& md, ws_state* st) {
  ws_model o{};
  o.n = md.n; o.m = md.m; o.p = md.p; o.K = md.K; o.C = md.C; o.S = md.S; o.n_point_local = md.nPointLocal; o.l_ref = md.lRef; o.par = md.par;
  o.node_offset = md.nodeOffset.data(); o.seg_offset = md.segOffset.data(); o.joint_offset = md.jointOffset.data(); o.strand_closed = md.strandClosed.data();
  o.seg_i0 = md.segI0.data(); o.seg_i1 = md.segI1.data(); o.j_seg0 = md.jSeg0.data(); o.j_seg1 = md.jSeg1.data(); o.j_node = md.jNode.data();
  o.l0 = md.l0.data(); o.j_len = md.jLen.data();
  o.EA = md.EA.data(); o.GA = md.GA.data(); o.EIx = md.EIx.data(); o.EIy = md.EIy.data(); o.GJ = md.GJ.data(); o.omega0 = md.omega0.data();
  o.c_seg = md.cSeg.data(); o.c_node = md.cNode.data(); o.c_len = md.cLen.data(); o.c_B = md.cB.data(); o.c_Q = md.cQ.data();
  o.mass = md.mass.data(); o.krot_static = md.krot.data(); o.load_const = md.loadConst.data();
  o.point_local_node = md.plNode.data(); o.point_local_F = md.plF.data(); o.node_free = md.nodeFree.data();
  o.k_strand = md.kStrand.data(); o.k_seg_off = md.kSegOff.data(); o.k_nseg = md.kNseg.data();
  o.k_wk = md.kwk.data(); o.k_hk = md.khk.data(); o.k_fc90 = md.kfc90.data(); o.k_E90 = md.kE90.data(); o.k_Hs = md.kHs.data();
  o.k_w = md.kw.data(); o.k_h = md.kh.data(); o.k_enabled = md.kEnabled.data(); o.k_locked = md.kLocked.data();
  o.node_w = md.nodeW.data(); o.node_h = md.nodeH.data();
  if (st) {
    *st = ws_state{};
    st->iteration = md.iteration; st->ke_prev = md.kePrev; st->residual = md.residual;
    st->X = md.X.data(); st->X0 = md.X0.data(); st->V = md.V.data(); st->Q = md.Q.data(); st->Q0 = md.Q0.data(); st->W = md.W.data();
    st->u = md.u.data(); st->n_prev = md.nPrev.data(); st->slip = md.slip.data(); st->stuck = md.stuck.data(); st->lam = md.lam.data();
    st->pen_acc = md.penAcc.data(); st->theta0 = md.theta0.data(); st->rot_slip = md.rotSlip.data(); st->dent = md.dent.data();
    st->climb = md.climb.data(); st->tanpsi = md.tanpsi.data(); st->dent_e = md.dentE.data(); st->floor_a = md.floorA.data();
    st->N = md.N.data(); st->Vsh = md.Vsh.data(); st->M = md.M.data(); st->cM = md.cM.data(); st->reactions = md.reactions.data();
    st->Fn = md.Fn.data(); st->Ft = md.Ft.data(); st->gap = md.gap.data(); st->kstate = md.kstate.data(); st->theta = md.theta.data();
    st->Anom = md.Anom.data(); st->pressure = md.pressure.data(); st->util = md.util.data(); st->Mrot = md.Mrot.data(); st->phi = md.phi.data();
    st->mid = md.mid.data(); st->ctype = md.ctype.data(); st->face_side = md.faceSide.data();
    st->floor_F = md.floorF.data(); st->floor_state = md.floorState.data();
  }
  return o;
}

}  // namespace

struct wsm_model { std::unique_ptr<Model> md; };

extern "C" {

WS_API const char* wsm_last_error(void) { return g_err.c_str(); }

WS_API void wsm_scene_defaults(wsm_scene* sc) {
  *sc = wsm_scene{};
  // MATERIAL / SECTION / BEAM (params.js:6-19), mm·N
  sc->grp = wsm_group{10000.0, 625.0, 700.0, 800.0, 8.0, 30.0, 60.0, 30.0, 0, 0, 0, 0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0};
  ws_params& p = sc->par;
  p.mass_scale = 1.5; p.damping = 1.0; p.tol = 1e-4; p.kinetic_damping = 1; p.warmup_iter = 300; p.max_iter = 5000;   // SOLVER
  p.mu = 0.35; p.strength = 1e4; p.stick_strength = 1e4; p.torsion_factor = 1.0; p.epsilon = 15.0; p.slip_rate = 1.0;  // CONTACT
  p.ramp_iter = 1500; p.rot_friction = 0; p.rot_factor = 1.0; p.seat_law = 0; p.seat_gap = 1; p.dent_rate = 0.01;
  p.dent_after = -1; p.edge_threshold = 0.5; p.augmented = 0; p.al_every = 600;
  p.floor_on = 0; p.floor_z = 0; p.floor_mu = 0.4;
  sc->contacts = 1; sc->prestress_asbuilt = 0;
  sc->subdivisions = 4; sc->overhang = 0.0; sc->offset_half_height = 1; sc->seed_strand_phase = 0;   // WEAVE
  sc->length_to_m = 1e-3; sc->force_to_n = 1.0;
  sc->gravity = 1; sc->gravity_dir[0] = 0; sc->gravity_dir[1] = 0; sc->gravity_dir[2] = -1;
  sc->udl = 0;
}

namespace {
int fromMesh(const double* V, int32_t nV, const int32_t* F, int32_t nF, int32_t do_weld, const wsm_scene* sc,
             const wsm_anchor* anchors, int32_t n_anchors, const wsm_point_load* loads, int32_t n_loads,
             const int32_t* flips, int32_t n_flips, const wsm_prep* prep, bool hybrid, const wsm_group* groups,
             int32_t n_groups, const int32_t* assign, int32_t n_assign, wsm_model** out,
             const int32_t* joint_strand = nullptr, const double* joint_s = nullptr, int32_t n_joints = 0,
             const int32_t* piece_group = nullptr, int32_t n_pieces = -1) {
  try {
    if (!V || !F || !sc || !out || nV <= 0 || nF <= 0) throw std::runtime_error("wsm_from_mesh: bad arguments");
    if (hybrid && ((n_groups > 0 && !groups) || (n_assign > 0 && !assign) || n_groups < 0 || n_assign < 0))
      throw std::runtime_error("wsm_from_mesh_groups: bad group arguments");
    if (n_pieces >= 0 && ((n_joints > 0 && (!joint_strand || !joint_s)) || (n_pieces > 0 && !piece_group) || n_joints < 0))
      throw std::runtime_error("wsm_from_mesh_pieces: bad joint or piece arguments");
    std::vector<double> Vv(V, V + 3 * size_t(nV)); std::vector<int32_t> Fv(F, F + 3 * size_t(nF));
    if (do_weld) weld(Vv, Fv);
    Mesh mesh = meshFromArrays(Vv, Fv);
    Topology topo; topo.mesh = &mesh;
    topo.subdivisions = sc->subdivisions; topo.overhang = sc->overhang; topo.halfHeight = sc->offset_half_height != 0;
    topo.seedPhase = sc->seed_strand_phase; topo.thickness = sc->grp.height;
    traceStrands(topo); assignParity(topo); applyFlips(topo, flips, n_flips);
    Setup su; su.sc = *sc;
    if (hybrid) {
      su.extra.assign(groups, groups + n_groups);
      su.strandGroup.assign(topo.nStrands(), 1);
      for (int a = 0; a < n_assign; a++) {
        const int s = assign[2 * a];
        if (s >= 0 && s < topo.nStrands()) su.strandGroup[s] = assign[2 * a + 1];
      }
      topo.strandH.resize(topo.nStrands());
      for (int s = 0; s < topo.nStrands(); s++) topo.strandH[s] = su.groupOfStrand(s).height;
    }
    if (n_pieces >= 0) {
      // Scene.hasPieceGroups (scene.js:71): the scene assigns pieces; joints of strands the topology does not have
      // are ignored, as jointPoints and the piece indices do (discretize.js:309, 337)
      su.pieceGroups = n_pieces > 0;
      su.joints.assign(topo.nStrands(), {});
      for (int q = 0; q < n_joints; q++) {
        const int s = joint_strand[q];
        if (s >= 0 && s < topo.nStrands()) su.joints[s].push_back(joint_s[q]);
      }
      for (auto& js : su.joints) std::sort(js.begin(), js.end());
      for (int q = 0; q < n_pieces; q++) su.pieces[fkey(piece_group[3 * q], piece_group[3 * q + 1])] = piece_group[3 * q + 2];
    }
    buildPolylines(topo); consistentFrames(topo);
    su.anchors.assign(anchors, anchors + n_anchors); su.loads.assign(loads, loads + n_loads);
    if (prep) { su.prep = *prep; su.hasPrep = true; }
    su.meshArea = meshArea(mesh);
    auto md = buildModel(fromTopology(topo), su);
    md->info.nV = mesh.nV; md->info.nF = mesh.nF; md->info.nE = mesh.nE; md->info.subdivisions = sc->subdivisions;
    md->info.parity_conflicts = topo.parityConflicts;
    *out = new wsm_model{std::move(md)};
    return WS_OK;
  } catch (const std::exception& e) { g_err = e.what(); return WS_ERR_ARG; }
}
}  // namespace

WS_API int wsm_from_mesh(const double* V, int32_t nV, const int32_t* F, int32_t nF, int32_t do_weld,
                         const wsm_scene* sc, const wsm_anchor* anchors, int32_t n_anchors,
                         const wsm_point_load* loads, int32_t n_loads, const int32_t* flips, int32_t n_flips,
                         const wsm_prep* prep, wsm_model** out) {
  return fromMesh(V, nV, F, nF, do_weld, sc, anchors, n_anchors, loads, n_loads, flips, n_flips, prep, false, nullptr,
                  0, nullptr, 0, out);
}

WS_API int wsm_from_mesh_groups(const double* V, int32_t nV, const int32_t* F, int32_t nF, int32_t do_weld,
                                const wsm_scene* sc, const wsm_anchor* anchors, int32_t n_anchors,
                                const wsm_point_load* loads, int32_t n_loads, const int32_t* flips, int32_t n_flips,
                                const wsm_prep* prep, const wsm_group* groups, int32_t n_groups,
                                const int32_t* strand_group, int32_t n_assign, wsm_model** out) {
  return fromMesh(V, nV, F, nF, do_weld, sc, anchors, n_anchors, loads, n_loads, flips, n_flips, prep, true, groups,
                  n_groups, strand_group, n_assign, out);
}

WS_API int wsm_from_mesh_pieces(const double* V, int32_t nV, const int32_t* F, int32_t nF, int32_t do_weld,
                                const wsm_scene* sc, const wsm_anchor* anchors, int32_t n_anchors,
                                const wsm_point_load* loads, int32_t n_loads, const int32_t* flips, int32_t n_flips,
                                const wsm_prep* prep, const wsm_group* groups, int32_t n_groups,
                                const int32_t* strand_group, int32_t n_assign, const int32_t* joint_strand,
                                const double* joint_s, int32_t n_joints, const int32_t* piece_group, int32_t n_pieces,
                                wsm_model** out) {
  if (n_pieces < 0) { g_err = "wsm_from_mesh_pieces: negative piece count"; return WS_ERR_ARG; }
  return fromMesh(V, nV, F, nF, do_weld, sc, anchors, n_anchors, loads, n_loads, flips, n_flips, prep, true, groups,
                  n_groups, strand_group, n_assign, out, joint_strand, joint_s, n_joints, piece_group, n_pieces);
}

WS_API int wsm_from_polylines(const double* P, const int32_t* counts, int32_t n_strands, const int32_t* knot_rows,
                              const int32_t* knot_counts, const double* normals, const int32_t* knots, int32_t n_knots,
                              const wsm_scene* sc, const wsm_anchor* anchors, int32_t n_anchors,
                              const wsm_point_load* loads, int32_t n_loads, wsm_model** out) {
  try {
    if (!P || !counts || n_strands <= 0 || !sc || !out) throw std::runtime_error("wsm_from_polylines: bad arguments");
    Built b; b.S = n_strands; b.overhang = sc->overhang;
    size_t off = 0, roff = 0;
    for (int s = 0; s < n_strands; s++) {
      int n = counts[s];
      std::vector<double> Pl(P + 3 * off, P + 3 * (off + n));
      std::vector<V3> hints(n, V3{0, 0, 1});
      if (normals) for (int i = 0; i < n; i++) hints[i] = V3{normals[3 * (off + i)], normals[3 * (off + i) + 1], normals[3 * (off + i) + 2]};
      b.nodeFrames.push_back(nodeFramesOf(Pl, hints, false));   // topology.js:348-363
      b.polylines.push_back(Pl); b.closed.push_back(0);
      int nk = knot_counts ? knot_counts[s] : 2;
      for (int k = 0; k < nk; k++) b.knotNode[fkey(s, k)] = knot_rows ? knot_rows[roff + k] : (k == 0 ? 0 : n - 1);
      b.nKnots.push_back(nk);
      roff += nk; off += n;
    }
    std::vector<std::pair<int, Built::KnotRec>> ks;
    // explicit crossings have no mesh edge: each is keyed by its index
    for (int c = 0; c < n_knots; c++) b.knots.push_back({knots[4 * c], knots[4 * c + 1], knots[4 * c + 2], knots[4 * c + 3], false, V3{0, 0, 1}, c, 0});
    Setup su; su.sc = *sc;
    su.anchors.assign(anchors, anchors + n_anchors); su.loads.assign(loads, loads + n_loads);
    auto md = buildModel(b, su);
    *out = new wsm_model{std::move(md)};
    return WS_OK;
  } catch (const std::exception& e) { g_err = e.what(); return WS_ERR_ARG; }
}

// Synthetic benchmark scene: a kagome weave on a hexagonal patch of a triangular lattice (edge a = 2·spacing/√3,
// R = round(span / (2·spacing)) rings, the lattice of make_meshes.py:13-29), raised onto a spherical cap whose central
// meridian has the developed length `span` over a plan width (1 − contraction)·span (sin θ / θ = 1 − contraction).
WS_API int wsm_hex_dome(const wsm_hex_spec* spec, const wsm_scene* sc_in, wsm_model** out) {
  try {
    if (!spec || !sc_in || !out) throw std::runtime_error("wsm_hex_dome: bad arguments");
    if (!(spec->span > 0 && spec->spacing > 0 && spec->segment > 0) || spec->contraction < 0 || spec->contraction >= 0.6)
      throw std::runtime_error("wsm_hex_dome: span, spacing, segment > 0 and 0 <= contraction < 0.6 required");
    wsm_scene sc = *sc_in;
    int R = std::max(1, int(std::lround(spec->span / (2.0 * spec->spacing))));
    double a = 2.0 * spec->spacing / std::sqrt(3.0);
    const double cfac = 1.0 - spec->contraction;
    // θ of the cap edge from sin θ / θ = 1 − c (bisection), cap radius from the developed half-width
    double th = 0.0;
    if (spec->contraction > 0) {
      double lo = 1e-9, hi = M_PI - 1e-9;
      for (int it = 0; it < 200; it++) { double mid = 0.5 * (lo + hi); if (std::sin(mid) / mid > cfac) lo = mid; else hi = mid; }
      th = 0.5 * (lo + hi);
    }
    double Rs = th > 0 ? 0.5 * spec->span / th : 0.0;       // cap radius: the central meridian keeps the developed half width
    std::vector<double> V; std::vector<int32_t> F;
    std::map<std::pair<int, int>, int> id;
    auto inHex = [&](int q, int r) { return std::abs(q) <= R && std::abs(r) <= R && std::abs(q + r) <= R; };
    for (int r = -R; r <= R; r++) for (int q = -R; q <= R; q++) {
      if (!inHex(q, r)) continue;
      double x = a * (q + 0.5 * r), y = a * r * std::sqrt(3.0) / 2.0;
      id[{q, r}] = int(V.size() / 3); V.push_back(x); V.push_back(y); V.push_back(0.0);
    }
    // radial-distance-preserving map of the flat lattice onto the cap: a vertex at flat radius rho goes to polar angle
    // rho / Rs (plan radius Rs sin(rho / Rs)); the flat-to-flat half width lands on plan radius (1 - c) span / 2
    if (Rs > 0) {
      double phiMax = 0.0;
      for (size_t i = 0; i < V.size() / 3; i++) phiMax = std::max(phiMax, std::hypot(V[3 * i], V[3 * i + 1]) / Rs);
      if (phiMax >= 0.5 * M_PI) throw std::runtime_error("wsm_hex_dome: contraction too large for the hexagon corners");
      for (size_t i = 0; i < V.size() / 3; i++) {
        double rho = std::hypot(V[3 * i], V[3 * i + 1]), ph = rho / Rs;
        double f = rho > 0 ? Rs * std::sin(ph) / rho : 1.0;
        V[3 * i] *= f; V[3 * i + 1] *= f; V[3 * i + 2] = Rs * (std::cos(ph) - std::cos(phiMax));
      }
    }
    for (int r = -R; r <= R; r++) for (int q = -R; q <= R; q++) {
      if (inHex(q, r) && inHex(q + 1, r) && inHex(q, r + 1)) { F.push_back(id[{q, r}]); F.push_back(id[{q + 1, r}]); F.push_back(id[{q, r + 1}]); }
      if (inHex(q + 1, r) && inHex(q + 1, r + 1) && inHex(q, r + 1)) { F.push_back(id[{q + 1, r}]); F.push_back(id[{q + 1, r + 1}]); F.push_back(id[{q, r + 1}]); }
    }
    sc.subdivisions = std::max(1, int(std::lround((a / 2.0) / spec->segment)));
    wsm_anchor an[2] = {{WSM_ALL_START, 0, 0, {1, 1, 1, 0, 0, 0}}, {WSM_ALL_END, 0, 0, {1, 1, 1, 0, 0, 0}}};
    wsm_prep prep{spec->noise, 2.0 * spec->spacing, spec->seed, 0};
    int rc = wsm_from_mesh(V.data(), int(V.size() / 3), F.data(), int(F.size() / 3), 1, &sc, an, 2, nullptr, 0, nullptr, 0, &prep, out);
    if (rc == WS_OK) {
      (*out)->md->info.rings = R;
      double zmax = 0; for (int i = 0; i < (*out)->md->n; i++) zmax = std::max(zmax, (*out)->md->X[3 * i + 2]);
      (*out)->md->info.rise = zmax;
    }
    return rc;
  } catch (const std::exception& e) { g_err = e.what(); return WS_ERR_ARG; }
}

WS_API int wsm_get(wsm_model* mdl, ws_model* model, ws_state* state) {
  if (!mdl) { g_err = "null model"; return WS_ERR_ARG; }
  ws_model o = describe(*mdl->md, state);
  if (model) *model = o;
  return WS_OK;
}

WS_API int wsm_info_get(wsm_model* mdl, wsm_info* info) { if (!mdl || !info) return WS_ERR_ARG; *info = mdl->md->info; return WS_OK; }

WS_API int wsm_segment_moments(const ws_model* md, const double* M, const double* cM, double* Mx, double* My, double* T) {
  if (!md || !Mx || !My || !T || (md->p && !M) || (md->C && !cM)) { g_err = "wsm_segment_moments: bad arguments"; return WS_ERR_ARG; }
  const int n = md->n, m = md->m;
  std::vector<double> nodeM(3 * size_t(n), 0.0), segM(3 * size_t(m), 0.0);
  std::vector<uint8_t> has0(n, 0);
  for (int j = 0; j < md->p; j++) {
    const int i = md->j_node[j];
    for (int k = 0; k < 3; k++) nodeM[3 * i + k] = M[3 * j + k];
    has0[i] = 1;
  }
  std::vector<uint8_t> hasB = has0;
  for (int c = 0; c < md->C; c++) {
    const int i = md->c_node[c];
    nodeM[3 * i] = cM[2 * c]; nodeM[3 * i + 1] = cM[2 * c + 1];
    hasB[i] = 1;
  }
  for (int j = 0; j < m; j++)
    for (int k = 0; k < 3; k++) segM[3 * j + k] = 0.5 * (nodeM[3 * md->seg_i0[j] + k] + nodeM[3 * md->seg_i1[j] + k]);
  // the bending pair falls back where an end has neither joint nor clamp, the torsion where it has no joint: first
  // over the start nodes, then over the end nodes (the later assignment wins when both ends lack one)
  for (int pass = 0; pass < 2; pass++) {
    const std::vector<uint8_t>& has = pass == 0 ? hasB : has0;
    const int k0 = pass == 0 ? 0 : 2, k1 = pass == 0 ? 2 : 3;
    for (int j = 0; j < m; j++) {
      const int a = md->seg_i0[j], b = md->seg_i1[j];
      for (int k = k0; k < k1; k++) if (!has[a]) segM[3 * j + k] = nodeM[3 * b + k];
    }
    for (int j = 0; j < m; j++) {
      const int a = md->seg_i0[j], b = md->seg_i1[j];
      for (int k = k0; k < k1; k++) if (!has[b]) segM[3 * j + k] = nodeM[3 * a + k];
    }
  }
  for (int j = 0; j < m; j++) { Mx[j] = segM[3 * j]; My[j] = segM[3 * j + 1]; T[j] = segM[3 * j + 2]; }
  return WS_OK;
}

// the stability check's DOFs, stencil and colouring (dofList, stencil, colour; stability.js:17-23, 57-108)
namespace {
struct StabLayout {
  std::vector<int32_t> kind, index, comp, cols, colour;
  std::vector<int64_t> rowPtr;
  int32_t ncol = 0;
};
StabLayout stabilityLayout(const ws_model& md, const double* u, int W) {
  StabLayout L;
  const int n = md.n, m = md.m, S = md.S;
  std::vector<std::vector<int32_t>> nodeDofs(n), segDofs(m);
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 3; k++)
      if (md.node_free[3 * i + k]) { nodeDofs[i].push_back(int32_t(L.kind.size())); L.kind.push_back(0); L.index.push_back(i); L.comp.push_back(k); }
  for (int j = 0; j < m; j++)
    for (int k = 0; k < 3; k++) { segDofs[j].push_back(int32_t(L.kind.size())); L.kind.push_back(1); L.index.push_back(j); L.comp.push_back(k); }
  const int nd = int(L.kind.size());
  std::vector<int32_t> nodeStrand(n), segStrand(m);
  for (int s = 0; s < S; s++) {
    for (int i = md.node_offset[s]; i < md.node_offset[s + 1]; i++) nodeStrand[i] = s;
    for (int j = md.seg_offset[s]; j < md.seg_offset[s + 1]; j++) segStrand[j] = s;
  }
  // the nodes and segments within w of position loc along lath s (modulo the lath when it is closed)
  auto window = [&](int s, int loc, int w, std::vector<int32_t>& out) {
    const int nn = md.node_offset[s + 1] - md.node_offset[s], ns = md.seg_offset[s + 1] - md.seg_offset[s];
    const bool closed = nn == ns;
    auto wrap = [](int v, int q) { int r = v % q; return r < 0 ? r + q : r; };
    for (int o = -w; o <= w; o++) {
      int a = loc + o;
      if (closed && nn) a = wrap(a, nn);
      if (a >= 0 && a < nn) for (int d : nodeDofs[md.node_offset[s] + a]) out.push_back(d);
      int b = loc + o;
      if (closed && ns) b = wrap(b, ns);
      if (b >= 0 && b < ns) for (int d : segDofs[md.seg_offset[s] + b]) out.push_back(d);
    }
  };
  std::vector<std::vector<int32_t>> N(nd);
  std::vector<int32_t> st;
  for (int d = 0; d < nd; d++) {
    const int s = L.kind[d] == 0 ? nodeStrand[L.index[d]] : segStrand[L.index[d]];
    const int loc = L.index[d] - (L.kind[d] == 0 ? md.node_offset[s] : md.seg_offset[s]);
    window(s, loc, W, N[d]);
  }
  if (md.par.contacts && md.K) {
    for (int c = 0; c < md.K; c++) {
      st.clear();
      window(md.k_strand[2 * c], int(std::floor(u[2 * c])), W + 2, st);
      window(md.k_strand[2 * c + 1], int(std::floor(u[2 * c + 1])), W + 2, st);
      for (int a : st) N[a].insert(N[a].end(), st.begin(), st.end());
    }
  }
  L.rowPtr.assign(nd + 1, 0);
  for (int d = 0; d < nd; d++) {
    std::sort(N[d].begin(), N[d].end());
    N[d].erase(std::unique(N[d].begin(), N[d].end()), N[d].end());
    L.rowPtr[d + 1] = L.rowPtr[d] + int64_t(N[d].size());
  }
  L.cols.reserve(size_t(L.rowPtr[nd]));
  for (int d = 0; d < nd; d++) L.cols.insert(L.cols.end(), N[d].begin(), N[d].end());
  // greedy distance-2 colouring in DOF order: the smallest colour no DOF sharing a row with d has taken.  Every row
  // keeps the set of colours taken in it (a bitset): d's forbidden colours are the union over its rows, since q
  // shares row r with d exactly when both lie in the (symmetric) stencil row r
  L.colour.assign(nd, -1);
  int words = 8;
  std::vector<uint64_t> taken(size_t(nd) * words, 0);
  std::vector<uint64_t> forbid;
  for (int d = 0; d < nd; d++) {
    forbid.assign(words, 0);
    for (int r : N[d]) {
      const uint64_t* t = &taken[size_t(r) * words];
      for (int w = 0; w < words; w++) forbid[w] |= t[w];
    }
    int c = -1;
    for (int w = 0; w < words && c < 0; w++)
      if (~forbid[w]) c = 64 * w + __builtin_ctzll(~forbid[w]);
    if (c < 0) {
      // every colour of the bitsets is taken: widen them
      c = 64 * words;
      std::vector<uint64_t> wider(size_t(nd) * 2 * words, 0);
      for (size_t r = 0; r < size_t(nd); r++) std::copy(&taken[r * words], &taken[r * words] + words, &wider[r * 2 * words]);
      taken.swap(wider);
      words *= 2;
    }
    L.colour[d] = c;
    L.ncol = std::max(L.ncol, c + 1);
    for (int r : N[d]) taken[size_t(r) * words + (c >> 6)] |= uint64_t(1) << (c & 63);
  }
  return L;
}
}  // namespace

WS_API int wsm_stability_layout(const ws_model* md, const double* u, int32_t W, int32_t* n_dofs, int64_t* nnz,
                                int32_t* n_colours, int32_t* dof_kind, int32_t* dof_index, int32_t* dof_comp,
                                int64_t* row_ptr, int32_t* cols, int32_t* colour) {
  try {
    if (!md || !n_dofs || !nnz || !n_colours || (md->par.contacts && md->K && !u) || W < 0)
      throw std::runtime_error("wsm_stability_layout: bad arguments");
    const StabLayout L = stabilityLayout(*md, u, W);
    const int nd = int(L.kind.size());
    if (dof_kind) {
      if (*n_dofs < nd || *nnz < L.rowPtr[nd] || !dof_index || !dof_comp || !row_ptr || !cols || !colour)
        throw std::runtime_error("wsm_stability_layout: output arrays smaller than the layout");
      std::copy(L.kind.begin(), L.kind.end(), dof_kind); std::copy(L.index.begin(), L.index.end(), dof_index);
      std::copy(L.comp.begin(), L.comp.end(), dof_comp); std::copy(L.rowPtr.begin(), L.rowPtr.end(), row_ptr);
      std::copy(L.cols.begin(), L.cols.end(), cols); std::copy(L.colour.begin(), L.colour.end(), colour);
    }
    *n_dofs = nd; *nnz = L.rowPtr[nd]; *n_colours = L.ncol;
    return WS_OK;
  } catch (const std::exception& e) { g_err = e.what(); return WS_ERR_ARG; }
}

WS_API int wsm_add_load(wsm_model* mdl, int32_t node, const double* F) {
  if (!mdl || !F || node < 0 || node >= mdl->md->n) { g_err = "wsm_add_load: bad node"; return WS_ERR_ARG; }
  for (int c = 0; c < 3; c++) mdl->md->loadConst[3 * node + c] += F[c];
  return WS_OK;
}

WS_API int wsm_set_strength(wsm_model* mdl, double strength, double stick) {
  if (!mdl) return WS_ERR_ARG;
  mdl->md->par.strength = strength; mdl->md->par.stick_strength = stick;
  return WS_OK;
}

WS_API int wsm_clone(wsm_model* mdl, wsm_model** out) {
  if (!mdl || !out) return WS_ERR_ARG;
  *out = new wsm_model{std::make_unique<Model>(*mdl->md)};
  return WS_OK;
}

WS_API int wsm_mesh_counts(wsm_model* mdl, int32_t* nV, int32_t* nF) {
  if (!mdl) return WS_ERR_ARG;
  if (nV) *nV = mdl->md->info.nV;
  if (nF) *nF = mdl->md->info.nF;
  return WS_OK;
}

WS_API int wsm_knots(wsm_model* mdl, int32_t* out, int32_t cap_rows, int32_t* K) {
  if (!mdl) return WS_ERR_ARG;
  const Model& md = *mdl->md;
  if (K) *K = md.K;
  if (out) std::copy(md.kKnot.begin(), md.kKnot.begin() + 6 * size_t(std::min(std::max(cap_rows, 0), md.K)), out);
  return WS_OK;
}

WS_API int wsm_frames(wsm_model* mdl, int32_t* out, int32_t cap_rows, int32_t* n_frames) {
  if (!mdl) return WS_ERR_ARG;
  const Model& md = *mdl->md;
  if (n_frames) *n_frames = int32_t(md.nodeFrameIds.size());
  if (!out) return WS_OK;
  int32_t r = 0;
  for (const auto& kv : md.nodeFrameIds) {
    if (r >= cap_rows) break;
    out[3 * r] = int32_t(kv.first / 1048576); out[3 * r + 1] = int32_t(kv.first % 1048576); out[3 * r + 2] = kv.second;
    r++;
  }
  return WS_OK;
}

// CpuSolver.lockContacts / unlockContacts (solver_cpu.js:869-871, PolylineContact.lock / unlock :141-142): the
// crossings on the listed mesh edges (all of them when edges is null); edges without a crossing are ignored
WS_API int wsm_lock_contacts(wsm_model* mdl, const int32_t* edges, int32_t n_edges, int32_t lock) {
  if (!mdl || n_edges < 0 || (n_edges > 0 && !edges)) return WS_ERR_ARG;
  Model& md = *mdl->md;
  std::set<int32_t> want(edges ? edges : nullptr, edges ? edges + n_edges : nullptr);
  for (int c = 0; c < md.K; c++)
    if (!edges || want.count(md.kKnot[6 * size_t(c)])) md.kLocked[c] = lock ? 1 : 0;
  return WS_OK;
}

WS_API void wsm_free(wsm_model* mdl) { delete mdl; }

WS_API int32_t wsm_abi_sizes(int64_t* out, int32_t cap) {
  const int64_t v[] = {sizeof(ws_params), sizeof(ws_model), sizeof(ws_state), sizeof(ws_control), sizeof(ws_batch_opts),
                       sizeof(ws_server_opts), sizeof(ws_result), sizeof(wsm_scene), sizeof(wsm_anchor), sizeof(wsm_point_load),
                       sizeof(wsm_prep), sizeof(wsm_hex_spec), sizeof(wsm_info), sizeof(ws_tune_entry)};
  int32_t n = int32_t(sizeof(v) / sizeof(v[0]));
  for (int32_t i = 0; i < n && i < cap; i++) out[i] = v[i];
  return n;
}

}  // extern "C"
