// CPU oracle for the polarizable-embedding field / induced-dipole path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: only tests/, the smoke check and the
// CPU-baseline leg of bench.py load it (through oracle/oracle.py). The product library never
// links, loads or calls it.
//
// It restates, in plain FP64 C++ (built with -ffp-contract=off), the algorithms of the reference
// (maxscheurer/cppe) for this path. Every function cites the reference file:line it follows.
// Operators of the fast multipole method are written from the Cartesian Taylor mathematics
// (generic multi-index loops), not from the reference's generated code.
//
// Parity pins (tests/test_oracle*.py):
//   direct fields m0/m1/m2      -> golden vector fmuls_loprop_solvated_20 (reference fixture)
//   damped kernels               -> sympy derivatives of the Thole potential f_V(aR)/R
//   B apply / CG / dense B       -> dense closed form inv(B) F on pna_6w + pna (reference test)
//   tree + lists + FMM operators -> ref_fmm_errors.csv columns (reference fixture), invariants
//   nuclear fields               -> closed form sum Z d/|d|^3
//   nuclear-multipole energy     -> golden -0.321349401430 (reference test)
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

// ---------------------------------------------------------------------------------------------
// Multi-index tables. Order-major; inside order k: x exponent descending, then y descending
// (xx, xy, xz, yy, yz, zz for k = 2), the component order of the reference's T tensors
// (src/core/tensors/tensors_recursive.cc:24-40 xyz2idx).
// ---------------------------------------------------------------------------------------------
inline int nterms(int p) { return p < 0 ? 0 : (p + 1) * (p + 2) * (p + 3) / 6; }  // fmm/utils.cc:5-11
inline int ncomp(int k) { return (k + 1) * (k + 2) / 2; }                            // math.cc:99

struct MIndex {
  int a, b, c;
};

struct MITable {
  int p = -1;
  std::vector<MIndex> mi;
  std::vector<int> lut;  // (a, b, c) -> flat index, size (p+2)^3
  std::vector<double> inv_fact;
  int dim = 0;
  void init(int pmax) {
    p   = pmax;
    dim = pmax + 2;
    mi.clear();
    lut.assign(dim * dim * dim, -1);
    for (int k = 0; k <= pmax; ++k)
      for (int a = k; a >= 0; --a)
        for (int b = k - a; b >= 0; --b) {
          int c = k - a - b;
          lut[(a * dim + b) * dim + c] = (int)mi.size();
          mi.push_back({a, b, c});
        }
    inv_fact.resize(mi.size());
    for (size_t i = 0; i < mi.size(); ++i) {
      double f = 1.0;
      for (int t = 2; t <= mi[i].a; ++t) f *= t;
      for (int t = 2; t <= mi[i].b; ++t) f *= t;
      for (int t = 2; t <= mi[i].c; ++t) f *= t;
      inv_fact[i] = 1.0 / f;
    }
  }
  int idx(int a, int b, int c) const {
    if (a < 0 || b < 0 || c < 0 || a + b + c > p) return -1;
    return lut[(a * dim + b) * dim + c];
  }
  int order(int i) const { return mi[i].a + mi[i].b + mi[i].c; }
};

// Derivatives D_n = d^n/dr^n (1/|r|) for |n| <= pmax by the Hermite-Coulomb recursion of
// McMurchie & Davidson (J. Comput. Phys. 26, 218 (1978)) in its point-charge limit:
//   R^{(j)}_{000} = (-1)^j (2j-1)!! / |r|^{2j+1},
//   R^{(j)}_{t+1,u,v} = t R^{(j+1)}_{t-1,u,v} + x R^{(j+1)}_{t,u,v}   (same for u, v).
void coulomb_derivs(const MITable& T, int pmax, double x, double y, double z, std::vector<double>& D) {
  const int n  = nterms(pmax);
  const double r2 = x * x + y * y + z * z;
  const double rinv = 1.0 / std::sqrt(r2);
  // aux[j][i] for i over multi-indices of order <= pmax - j
  std::vector<std::vector<double>> aux(pmax + 1, std::vector<double>(n, 0.0));
  double base = rinv;  // j = 0
  double dfac = 1.0;   // (2j-1)!!
  for (int j = 0; j <= pmax; ++j) {
    double sgn = (j % 2) ? -1.0 : 1.0;
    aux[j][0] = sgn * dfac * base;
    base *= rinv * rinv;
    dfac *= (2 * j + 1);
  }
  for (int k = 1; k <= pmax; ++k) {
    for (int j = 0; j + k <= pmax; ++j) {
      for (int i = nterms(k - 1); i < nterms(k); ++i) {
        const MIndex m = T.mi[i];
        double v;
        if (m.a > 0) {
          int i1 = T.idx(m.a - 1, m.b, m.c), i2 = T.idx(m.a - 2, m.b, m.c);
          v = x * aux[j + 1][i1] + (m.a - 1 > 0 ? (m.a - 1) * aux[j + 1][i2] : 0.0);
        } else if (m.b > 0) {
          int i1 = T.idx(m.a, m.b - 1, m.c), i2 = T.idx(m.a, m.b - 2, m.c);
          v = y * aux[j + 1][i1] + (m.b - 1 > 0 ? (m.b - 1) * aux[j + 1][i2] : 0.0);
        } else {
          int i1 = T.idx(m.a, m.b, m.c - 1), i2 = T.idx(m.a, m.b, m.c - 2);
          v = z * aux[j + 1][i1] + (m.c - 1 > 0 ? (m.c - 1) * aux[j + 1][i2] : 0.0);
        }
        aux[j][i] = v;
      }
    }
  }
  D.assign(aux[0].begin(), aux[0].end());
}

// ---------------------------------------------------------------------------------------------
// Interaction tensors of rank 1..3, undamped or Thole-damped.
// Thole screening factors: src/core/tensors/tensors_recursive.cc:93-114; applied per term of
// the Cartesian tensor as in T() (tensors_recursive.cc:42-91): rank-1 term x f_E; rank-2
// delta term x f_E, dd term x f_T; rank-3 d*delta terms x f_T, ddd term x f_D.
// ---------------------------------------------------------------------------------------------
struct Screen {
  double fE = 1.0, fT = 1.0, fD = 1.0, fV = 1.0;
};

Screen thole_factors(double v) {
  Screen s;
  const double e = std::exp(-v);
  s.fV = 1.0 - (0.5 * v + 1.0) * e;
  s.fE = s.fV - (0.5 * v * v + 0.5 * v) * e;
  s.fT = s.fE - (1.0 / 6.0) * v * v * v * e;
  s.fD = s.fT - (1.0 / 30.0) * v * v * v * v * e;
  return s;
}

// Field at a target from one source of order m at displacement d = r_target - r_source,
// accumulated into F. Convention of the reference's P2P<m,3> kernels (fmm/field_m{0,1,2}.cc):
// F = -grad phi with phi = sum_j S_j d^j G(d), G = 1/R (or f_V(aR)/R when damped), i.e.
//   m=0: +q d/R^3;  m=1: -T2 mu (negative dipole field);  m=2: +field of the half-diagonal
// packed quadrupole (S = 1/2 Txx, Txy, Txz, 1/2 Tyy, Tyz, 1/2 Tzz).
void pair_field(int m, double dx, double dy, double dz, const double* S, double a, bool damped,
                double* F) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double R  = std::sqrt(r2);
  Screen s;
  if (damped) s = thole_factors(a * R);
  const double ir  = 1.0 / R;
  const double ir2 = ir * ir;
  const double ir3 = ir * ir2;
  const double ir5 = ir3 * ir2;
  const double ir7 = ir5 * ir2;
  const double d[3] = {dx, dy, dz};
  if (m == 0) {
    // T1_a = -f_E d_a / R^3 ; F = -q T1
    for (int a = 0; a < 3; ++a) F[a] += S[0] * s.fE * d[a] * ir3;
  } else if (m == 1) {
    // T2_ab = 3 f_T d_a d_b / R^5 - f_E delta_ab / R^3 ; F = -T2 S
    double dS = d[0] * S[0] + d[1] * S[1] + d[2] * S[2];
    for (int a = 0; a < 3; ++a) F[a] += -(3.0 * s.fT * d[a] * dS * ir5 - s.fE * S[a] * ir3);
  } else {
    // T3_abc = -15 f_D d_a d_b d_c / R^7 + 3 f_T (d_a delta_bc + d_b delta_ac + d_c delta_ab)/R^5
    // F_a = -sum_{|j|=2} S_j T3[a + j]; j runs over xx, xy, xz, yy, yz, zz.
    static const int jb[6] = {0, 0, 0, 1, 1, 2};
    static const int jc[6] = {0, 1, 2, 1, 2, 2};
    for (int a = 0; a < 3; ++a) {
      double acc = 0.0;
      for (int q = 0; q < 6; ++q) {
        int b = jb[q], c = jc[q];
        double t3 = -15.0 * s.fD * d[a] * d[b] * d[c] * ir7 +
                    3.0 * s.fT *
                          (d[a] * (b == c ? 1.0 : 0.0) + d[b] * (a == c ? 1.0 : 0.0) +
                           d[c] * (a == b ? 1.0 : 0.0)) *
                          ir5;
        acc += S[q] * t3;
      }
      F[a] += -acc;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Particles (fmm/tree.hh:16-27) and the direct sums (fmm/calculate.cc:262-300).
// ---------------------------------------------------------------------------------------------
struct Particles {
  int n = 0;
  int m = 0;                 // source order
  const double* pos = nullptr;  // AoS3
  const double* S = nullptr;    // n * ncomp(m)
  const double* alpha = nullptr;  // isotropic polarizability, 0 = not damped
  const int* excl_off = nullptr;  // CSR of exclusions (values compared to particle positions)
  const int* excl_idx = nullptr;
  bool excludes(int i, int j) const {  // fmm/tree.hh:23-26 (std::find over the list)
    if (!excl_off) return false;
    for (int k = excl_off[i]; k < excl_off[i + 1]; ++k)
      if (excl_idx[k] == j) return true;
    return false;
  }
};

void direct_field(const Particles& P, double damping, double* F) {
  const int cs = ncomp(P.m);
  // calculate.cc:262-300; F is accumulated, not zeroed (callers pass zeroed buffers)
#pragma omp parallel for schedule(runtime)
  for (int i = 0; i < P.n; ++i) {
    for (int j = 0; j < P.n; ++j) {
      if (i == j || P.excludes(i, j)) continue;
      double dx = P.pos[3 * i] - P.pos[3 * j];
      double dy = P.pos[3 * i + 1] - P.pos[3 * j + 1];
      double dz = P.pos[3 * i + 2] - P.pos[3 * j + 2];
      bool damped = false;
      double v    = 0.0;
      if (damping > 0.0 && P.alpha) {
        double a1 = P.alpha[i], a2 = P.alpha[j];
        if (std::abs(a1) > 0.0 && std::abs(a2) > 0.0) {
          v      = damping / std::pow(a1 * a2, 1.0 / 6.0);
          damped = true;
        }
      }
      pair_field(P.m, dx, dy, dz, &P.S[cs * j], v, damped, &F[3 * i]);
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Octree: fmm/tree.hh:29-89 (Cell), tree.cc:94-132 (add_child, split_cell),
// tree.cc:134-237 (build_shared_tree).
// ---------------------------------------------------------------------------------------------
struct Cell {
  size_t nleaf = 0;
  int nchild   = 0;  // bit mask of occupied octants
  size_t level = 0;
  size_t child[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<size_t> leaf;
  double x, y, z, r, rmax;
  size_t parent;
};

struct Tree {
  int ncrit = 64, order = 5;
  double theta = 0.5;
  std::vector<Cell> cells;
  std::vector<std::pair<size_t, size_t>> m2l;  // (source B, target A), fmm/calculate.cc:87
  std::vector<std::pair<size_t, size_t>> p2p;  // (target A, source B), calculate.cc:99
};

Cell make_cell(double x, double y, double z, double r, size_t parent, size_t level, size_t ncrit) {
  Cell c;  // tree.cc:11-24
  c.x = x;
  c.y = y;
  c.z = z;
  c.r = r;
  c.rmax   = std::sqrt(3 * r * r);
  c.parent = parent;
  c.level  = level;
  c.leaf.assign(ncrit, 0);
  return c;
}

void add_child(std::vector<Cell>& cells, int octant, size_t p, size_t ncrit) {  // tree.cc:94-109
  size_t c = cells.size();
  double r = cells[p].r / 2.0;
  double x = cells[p].x + r * ((octant & 1) * 2 - 1);
  double y = cells[p].y + r * ((octant & 2) - 1);
  double z = cells[p].z + r * ((octant & 4) / 2 - 1);
  cells.push_back(make_cell(x, y, z, r, p, cells[p].level + 1, ncrit));
  cells[p].child[octant] = c;
  cells[c].nleaf         = 0;
  cells[p].nchild |= (1 << octant);
}

void split_cell(std::vector<Cell>& cells, const double* pos, size_t p, size_t ncrit) {  // tree.cc:111-132
  for (size_t i = 0; i < cells[p].leaf.size(); ++i) {
    size_t l   = cells[p].leaf[i];
    int octant = (pos[3 * l] > cells[p].x) + ((pos[3 * l + 1] > cells[p].y) << 1) +
                 ((pos[3 * l + 2] > cells[p].z) << 2);
    if (!(cells[p].nchild & (1 << octant))) add_child(cells, octant, p, ncrit);
    size_t c                      = cells[p].child[octant];
    cells[c].leaf[cells[c].nleaf] = l;
    cells[c].nleaf += 1;
    if (cells[c].nleaf >= ncrit) split_cell(cells, pos, c, ncrit);
  }
}

void build_tree(Tree& T, int n, const double* pos, int ncrit) {  // tree.cc:134-237
  T.cells.clear();
  T.ncrit = ncrit;
  double xavg = 0, yavg = 0, zavg = 0;
  for (int i = 0; i < n; ++i) {
    xavg += pos[3 * i];
    yavg += pos[3 * i + 1];
    zavg += pos[3 * i + 2];
  }
  xavg /= n;
  yavg /= n;
  zavg /= n;
  double xmax = 0, ymax = 0, zmax = 0;
  for (int i = 0; i < n; ++i) {
    double x = std::abs(pos[3 * i] - xavg);
    double y = std::abs(pos[3 * i + 1] - yavg);
    double z = std::abs(pos[3 * i + 2] - zavg);
    if (x > xmax) xmax = x;
    if (y > ymax) ymax = y;
    if (z > zmax) zmax = z;
  }
  double r = (xmax > ymax ? (xmax > zmax ? xmax : zmax) : (ymax > zmax ? ymax : zmax)) * 1.001;
  T.cells.push_back(make_cell(xavg, yavg, zavg, r, 0, 0, ncrit));
  for (int i = 0; i < n; ++i) {
    size_t curr = 0;
    while (T.cells[curr].nleaf >= (size_t)ncrit) {
      T.cells[curr].nleaf += 1;
      int octant = (pos[3 * i] > T.cells[curr].x) + ((pos[3 * i + 1] > T.cells[curr].y) << 1) +
                   ((pos[3 * i + 2] > T.cells[curr].z) << 2);
      if (!(T.cells[curr].nchild & (1 << octant))) add_child(T.cells, octant, curr, ncrit);
      curr = T.cells[curr].child[octant];
    }
    T.cells[curr].leaf[T.cells[curr].nleaf] = i;
    T.cells[curr].nleaf += 1;
    if (T.cells[curr].nleaf >= (size_t)ncrit) split_cell(T.cells, pos, curr, ncrit);
  }
}

// Dual-tree traversal, fmm/calculate.cc:74-127.
void interact(size_t A, size_t B, Tree& T) {
  const auto& cells = T.cells;
  const double dx = cells[A].x - cells[B].x;
  const double dy = cells[A].y - cells[B].y;
  const double dz = cells[A].z - cells[B].z;
  const double R  = std::sqrt(dx * dx + dy * dy + dz * dz);
  if (R * T.theta > (cells[A].rmax + cells[B].rmax)) {
    T.m2l.push_back({B, A});
  } else if (cells[A].nchild == 0 && cells[B].nchild == 0) {
    if (cells[B].nleaf >= (size_t)T.ncrit) {
      // unreachable: a cell holding >= ncrit particles always has children (calculate.cc:93-97)
      T.m2l.push_back({B, A});
    } else {
      T.p2p.push_back({A, B});
    }
  } else if (cells[B].nchild == 0 || (cells[A].rmax >= cells[B].rmax && cells[A].nchild != 0)) {
    for (int oa = 0; oa < 8; ++oa)
      if (cells[A].nchild & (1 << oa)) interact(cells[A].child[oa], B, T);
  } else {
    for (int ob = 0; ob < 8; ++ob)
      if (cells[B].nchild & (1 << ob)) interact(A, cells[B].child[ob], T);
  }
}

void build_lists(Tree& T, double theta) {
  T.theta = theta;
  T.m2l.clear();
  T.p2p.clear();
  interact(0, 0, T);
  // tree.cc:224-227 sorts by source cell (unstable); a stable sort keeps this oracle deterministic
  std::stable_sort(T.m2l.begin(), T.m2l.end(),
                   [](const std::pair<size_t, size_t>& l, const std::pair<size_t, size_t>& r) {
                     return l.first < r.first;
                   });
}

// ---------------------------------------------------------------------------------------------
// FMM passes (fmm/tree.cc:359-385, fmm/calculate.cc:129-260) with Cartesian Taylor operators
// truncated at |n| + |m| <= p. M holds moments of order m..p, L Taylor coefficients of order
// 0..p-m (fmm/utils.cc:13-20 Msize/Lsize); both are stored here in full multi-index arrays.
//   potential  phi(x) = sum_j M_j(z) D_j(x - z),   M_j(z) = sum_b (z - x_b)^j / j! (x) source
//   local      phi(x) = sum_n (x - z_A)^n / n! L_n, L_n = sum_j M_j D_{n+j}(z_A - z_B)
//   field      F = -grad phi
// ---------------------------------------------------------------------------------------------
struct Powers {
  double px[16], py[16], pz[16];
  void set(int p, double x, double y, double z) {
    px[0] = py[0] = pz[0] = 1.0;
    for (int k = 1; k <= p; ++k) {
      px[k] = px[k - 1] * x;
      py[k] = py[k - 1] * y;
      pz[k] = pz[k - 1] * z;
    }
  }
  double mono(const MIndex& m) const { return px[m.a] * py[m.b] * pz[m.c]; }
};

// M_parent += shift(M_child) with d = z_parent - z_child (calculate.cc:159-184; P2M is the same
// shift applied to a point multipole, calculate.cc:129-157).
void m2m(const MITable& T, int p, int msrc, const double* Mc, double* Mp, double dx, double dy,
         double dz) {
  Powers pw;
  pw.set(p, dx, dy, dz);
  for (int n = nterms(msrc - 1); n < nterms(p); ++n) {
    const MIndex mn = T.mi[n];
    double acc      = 0.0;
    for (int ka = 0; ka <= mn.a; ++ka)
      for (int kb = 0; kb <= mn.b; ++kb)
        for (int kc = 0; kc <= mn.c; ++kc) {
          int j = T.idx(mn.a - ka, mn.b - kb, mn.c - kc);
          if (T.order(j) < msrc) continue;
          int k = T.idx(ka, kb, kc);
          acc += pw.mono(T.mi[k]) * T.inv_fact[k] * Mc[j];
        }
    Mp[n] += acc;
  }
}

void m2l(const MITable& T, int p, int msrc, const double* M, double* L, double dx, double dy,
         double dz, std::vector<double>& D) {
  coulomb_derivs(T, p, dx, dy, dz, D);
  for (int n = 0; n < nterms(p - msrc); ++n) {
    const MIndex mn = T.mi[n];
    int on          = T.order(n);
    double acc      = 0.0;
    for (int j = nterms(msrc - 1); j < nterms(p - on); ++j) {
      const MIndex mj = T.mi[j];
      acc += M[j] * D[T.idx(mn.a + mj.a, mn.b + mj.b, mn.c + mj.c)];
    }
    L[n] += acc;
  }
}

// L_child += shift(L_parent), d = z_child - z_parent (calculate.cc:227-243)
void l2l(const MITable& T, int lord, const double* Lp, double* Lc, double dx, double dy, double dz) {
  Powers pw;
  pw.set(lord, dx, dy, dz);
  for (int n = 0; n < nterms(lord); ++n) {
    const MIndex mn = T.mi[n];
    double acc      = 0.0;
    for (int k = 0; k < nterms(lord - T.order(n)); ++k) {
      const MIndex mk = T.mi[k];
      acc += pw.mono(mk) * T.inv_fact[k] * Lp[T.idx(mn.a + mk.a, mn.b + mk.b, mn.c + mk.c)];
    }
    Lc[n] += acc;
  }
}

// F += -grad of the local expansion at h = x - z_cell (calculate.cc:245-260)
void l2p(const MITable& T, int lord, const double* L, double* F, double hx, double hy, double hz) {
  Powers pw;
  pw.set(lord, hx, hy, hz);
  for (int a = 0; a < 3; ++a) {
    double acc = 0.0;
    for (int k = 0; k < nterms(lord - 1); ++k) {
      const MIndex mk = T.mi[k];
      int t           = T.idx(mk.a + (a == 0), mk.b + (a == 1), mk.c + (a == 2));
      acc += pw.mono(mk) * T.inv_fact[k] * L[t];
    }
    F[a] -= acc;
  }
}

void p2p_cells(const Tree& T, const Particles& P, size_t A, size_t B, double damping, double* F) {
  // calculate.cc:29-72
  const int cs = ncomp(P.m);
  for (size_t p1 = 0; p1 < T.cells[A].nleaf; ++p1) {
    size_t l1 = T.cells[A].leaf[p1];
    for (size_t p2 = 0; p2 < T.cells[B].nleaf; ++p2) {
      size_t l2 = T.cells[B].leaf[p2];
      if (l2 == l1 || P.excludes((int)l1, (int)l2)) continue;
      double dx = P.pos[3 * l1] - P.pos[3 * l2];
      double dy = P.pos[3 * l1 + 1] - P.pos[3 * l2 + 1];
      double dz = P.pos[3 * l1 + 2] - P.pos[3 * l2 + 2];
      bool damped = false;
      double v    = 0.0;
      if (damping > 0.0 && P.alpha) {
        double a1 = P.alpha[l1], a2 = P.alpha[l2];
        if (a1 > 0.0 && a2 > 0.0) {
          v      = damping / std::pow(a1 * a2, 1.0 / 6.0);
          damped = true;
        }
      }
      pair_field(P.m, dx, dy, dz, &P.S[cs * l2], v, damped, &F[3 * l1]);
    }
  }
}

void fmm_field(const Tree& T, const Particles& P, double damping, double* F) {
  const int p    = T.order;
  const int msrc = P.m;
  const int cs   = ncomp(msrc);
  MITable MT;
  MT.init(p + 1);
  const size_t nc = T.cells.size();
  const int nt    = nterms(p);
  std::vector<double> M(nc * nt, 0.0), L(nc * nt, 0.0);
  for (int i = 0; i < 3 * P.n; ++i) F[i] = 0.0;  // tree.cc:362-364
  // P2M (calculate.cc:129-157): each particle is a point multipole of order msrc.
#pragma omp parallel for schedule(dynamic, 16)
  for (size_t c = 0; c < nc; ++c) {
    if (T.cells[c].nleaf >= (size_t)T.ncrit) continue;
    std::vector<double> Q(nt, 0.0);
    for (size_t i = 0; i < T.cells[c].nleaf; ++i) {
      size_t l = T.cells[c].leaf[i];
      std::fill(Q.begin(), Q.end(), 0.0);
      for (int q = 0; q < cs; ++q) Q[nterms(msrc - 1) + q] = P.S[cs * l + q];
      m2m(MT, p, msrc, Q.data(), &M[c * nt], T.cells[c].x - P.pos[3 * l],
          T.cells[c].y - P.pos[3 * l + 1], T.cells[c].z - P.pos[3 * l + 2]);
    }
  }
  // M2M, children to parents in reverse creation order (calculate.cc:177-183)
  for (size_t c = nc - 1; c > 0; --c) {
    size_t pa = T.cells[c].parent;
    m2m(MT, p, msrc, &M[c * nt], &M[pa * nt], T.cells[pa].x - T.cells[c].x,
        T.cells[pa].y - T.cells[c].y, T.cells[pa].z - T.cells[c].z);
  }
  // M2L grouped by target so that threads own their L (the reference uses atomics instead)
  std::vector<std::vector<size_t>> by_target(nc);
  for (auto& pr : T.m2l) by_target[pr.second].push_back(pr.first);
#pragma omp parallel
  {
    std::vector<double> D;
#pragma omp for schedule(dynamic, 16)
    for (size_t A = 0; A < nc; ++A)
      for (size_t B : by_target[A])
        m2l(MT, p, msrc, &M[B * nt], &L[A * nt], T.cells[A].x - T.cells[B].x,
            T.cells[A].y - T.cells[B].y, T.cells[A].z - T.cells[B].z, D);
  }
  // P2P lists grouped by target leaf (calculate.cc:205-225)
  std::vector<std::vector<size_t>> p2p_by_target(nc);
  for (auto& pr : T.p2p) p2p_by_target[pr.first].push_back(pr.second);
#pragma omp parallel for schedule(dynamic, 16)
  for (size_t A = 0; A < nc; ++A)
    for (size_t B : p2p_by_target[A]) p2p_cells(T, P, A, B, damping, F);
  // L2L in creation order (calculate.cc:227-243)
  const int lord = p - msrc;
  for (size_t pa = 0; pa < nc; ++pa)
    for (int o = 0; o < 8; ++o)
      if (T.cells[pa].nchild & (1 << o)) {
        size_t c = T.cells[pa].child[o];
        l2l(MT, lord, &L[pa * nt], &L[c * nt], T.cells[c].x - T.cells[pa].x,
            T.cells[c].y - T.cells[pa].y, T.cells[c].z - T.cells[pa].z);
      }
  // L2P on leaves (calculate.cc:245-260)
#pragma omp parallel for schedule(dynamic, 16)
  for (size_t c = 0; c < nc; ++c) {
    if (T.cells[c].nleaf >= (size_t)T.ncrit) continue;
    for (size_t i = 0; i < T.cells[c].nleaf; ++i) {
      size_t k = T.cells[c].leaf[i];
      l2p(MT, lord, &L[c * nt], &F[3 * k], P.pos[3 * k] - T.cells[c].x,
          P.pos[3 * k + 1] - T.cells[c].y, P.pos[3 * k + 2] - T.cells[c].z);
    }
  }
}

// 3x3 helpers (packed xx, xy, xz, yy, yz, zz; math.cc:7-31)
void unpack(const double* p6, double m[3][3]) {
  m[0][0] = p6[0];
  m[0][1] = m[1][0] = p6[1];
  m[0][2] = m[2][0] = p6[2];
  m[1][1] = p6[3];
  m[1][2] = m[2][1] = p6[4];
  m[2][2] = p6[5];
}

void inverse3(const double a[3][3], double inv[3][3]) {
  double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
  double c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
  double c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
  double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
  double id  = 1.0 / det;
  inv[0][0]  = c00 * id;
  inv[1][0]  = c01 * id;
  inv[2][0]  = c02 * id;
  inv[0][1]  = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
  inv[1][1]  = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
  inv[2][1]  = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
  inv[0][2]  = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
  inv[1][2]  = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
  inv[2][2]  = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
}

// ---------------------------------------------------------------------------------------------
// The classical response operator and its CG solver (bmatrix.{hh,cc}, electric_fields.cc:191-239).
// ---------------------------------------------------------------------------------------------
struct BSystem {
  int n = 0;
  const double* pos = nullptr;
  const double* alpha6 = nullptr;  // packed polarizabilities
  std::vector<double> alpha_inv;   // 9 per site
  std::vector<double> alpha_iso;   // trace/3, potential.hh:68
  const int* excl_off = nullptr;
  const int* excl_idx = nullptr;
  double damping = 0.0;  // 0 = undamped (bmatrix.cc:18-21)
  int fmm = 0;
  int ncrit = 64, order = 5;
  double theta = 0.5;
  int rebuild_per_apply = 0;  // the reference rebuilds the tree in every apply (bmatrix.cc:28-29)
  Tree tree;
  bool have_tree = false;
};

void bsys_init(BSystem& B) {
  B.alpha_inv.resize(9 * B.n);
  B.alpha_iso.resize(B.n);
  for (int i = 0; i < B.n; ++i) {
    double a[3][3], ai[3][3];
    unpack(&B.alpha6[6 * i], a);
    inverse3(a, ai);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) B.alpha_inv[9 * i + 3 * r + c] = ai[r][c];
    B.alpha_iso[i] = (a[0][0] + a[1][1] + a[2][2]) / 3.0;
  }
}

void bsys_apply(BSystem& B, const double* x, double* out) {  // bmatrix.cc:9-40
  std::vector<double> F(3 * B.n, 0.0);
  Particles P;
  P.n        = B.n;
  P.m        = 1;
  P.pos      = B.pos;
  P.S        = x;
  P.alpha    = B.damping > 0.0 ? B.alpha_iso.data() : nullptr;
  P.excl_off = B.excl_off;
  P.excl_idx = B.excl_idx;
  if (B.fmm) {
    if (!B.have_tree || B.rebuild_per_apply) {
      B.tree = Tree();
      B.tree.order = B.order;
      build_tree(B.tree, B.n, B.pos, B.ncrit);
      build_lists(B.tree, B.theta);
      B.have_tree = true;
    }
    fmm_field(B.tree, P, B.damping, F.data());
  } else {
    direct_field(P, B.damping, F.data());
  }
#pragma omp parallel for
  for (int i = 0; i < B.n; ++i)
    for (int r = 0; r < 3; ++r) {
      const double* ai = &B.alpha_inv[9 * i + 3 * r];
      out[3 * i + r]   = F[3 * i + r] + (ai[0] * x[3 * i] + ai[1] * x[3 * i + 1] + ai[2] * x[3 * i + 2]);
    }
}

void apply_alpha(const BSystem& B, const double* x, double* out) {  // bmatrix.cc:42-51
#pragma omp parallel for
  for (int i = 0; i < B.n; ++i) {
    double a[3][3];
    unpack(&B.alpha6[6 * i], a);
    for (int r = 0; r < 3; ++r)
      out[3 * i + r] = a[r][0] * x[3 * i] + a[r][1] * x[3 * i + 1] + a[r][2] * x[3 * i + 2];
  }
}

double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Field of order-m sources at every particle, direct summation (calculate.cc:262-300).
// alpha_iso may be NULL (no damping). F is accumulated.
void orc_direct_field(int m, int n, const double* pos, const double* S, const double* alpha_iso,
                      const int* excl_off, const int* excl_idx, double damping, double* F) {
  Particles P;
  P.n = n;
  P.m = m;
  P.pos = pos;
  P.S = S;
  P.alpha = alpha_iso;
  P.excl_off = excl_off;
  P.excl_idx = excl_idx;
  direct_field(P, damping, F);
}

// Field at one list of targets from all particles (brute force for sampled error checks).
void orc_direct_field_targets(int m, int n, const double* pos, const double* S,
                              const double* alpha_iso, const int* excl_off, const int* excl_idx,
                              double damping, int nt, const int* targets, double* F) {
  Particles P;
  P.n = n;
  P.m = m;
  P.pos = pos;
  P.S = S;
  P.alpha = alpha_iso;
  P.excl_off = excl_off;
  P.excl_idx = excl_idx;
  const int cs = ncomp(m);
#pragma omp parallel for schedule(dynamic, 4)
  for (int t = 0; t < nt; ++t) {
    int i = targets[t];
    double acc[3] = {0, 0, 0};
    for (int j = 0; j < n; ++j) {
      if (i == j || P.excludes(i, j)) continue;
      bool damped = false;
      double v = 0.0;
      if (damping > 0.0 && alpha_iso) {
        double a1 = alpha_iso[i], a2 = alpha_iso[j];
        if (std::abs(a1) > 0.0 && std::abs(a2) > 0.0) {
          v = damping / std::pow(a1 * a2, 1.0 / 6.0);
          damped = true;
        }
      }
      pair_field(m, pos[3 * i] - pos[3 * j], pos[3 * i + 1] - pos[3 * j + 1],
                 pos[3 * i + 2] - pos[3 * j + 2], &S[cs * j], v, damped, acc);
    }
    F[3 * t] = acc[0];
    F[3 * t + 1] = acc[1];
    F[3 * t + 2] = acc[2];
  }
}

// Build the octree + interaction lists. Returns an opaque tree, or NULL on error.
void* orc_tree_build(int n, const double* pos, int ncrit, int order, double theta) {
  try {
    Tree* T  = new Tree();
    T->order = order;
    build_tree(*T, n, pos, ncrit);
    build_lists(*T, theta);
    return T;
  } catch (std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

void orc_tree_free(void* t) { delete static_cast<Tree*>(t); }

// sizes: [ncells, n_m2l, n_p2p]
void orc_tree_sizes(void* t, long long* out) {
  Tree* T = static_cast<Tree*>(t);
  out[0]  = (long long)T->cells.size();
  out[1]  = (long long)T->m2l.size();
  out[2]  = (long long)T->p2p.size();
}

// Per cell: geo[5*c] = x, y, z, r, rmax; meta[4*c] = parent, level, nchild mask, nleaf;
// leafs[(ncrit)*c + i] for i < min(nleaf, ncrit-1) when a leaf (else -1).
void orc_tree_cells(void* t, double* geo, long long* meta, long long* leafs) {
  Tree* T = static_cast<Tree*>(t);
  for (size_t c = 0; c < T->cells.size(); ++c) {
    const Cell& C = T->cells[c];
    geo[5 * c] = C.x;
    geo[5 * c + 1] = C.y;
    geo[5 * c + 2] = C.z;
    geo[5 * c + 3] = C.r;
    geo[5 * c + 4] = C.rmax;
    meta[4 * c] = (long long)C.parent;
    meta[4 * c + 1] = (long long)C.level;
    meta[4 * c + 2] = C.nchild;
    meta[4 * c + 3] = (long long)C.nleaf;
    bool is_leaf = C.nchild == 0;
    for (int i = 0; i < T->ncrit; ++i)
      leafs[(size_t)T->ncrit * c + i] = (is_leaf && (size_t)i < C.nleaf) ? (long long)C.leaf[i] : -1;
  }
}

void orc_tree_lists(void* t, long long* m2l, long long* p2p) {
  Tree* T = static_cast<Tree*>(t);
  for (size_t i = 0; i < T->m2l.size(); ++i) {
    m2l[2 * i] = (long long)T->m2l[i].first;
    m2l[2 * i + 1] = (long long)T->m2l[i].second;
  }
  for (size_t i = 0; i < T->p2p.size(); ++i) {
    p2p[2 * i] = (long long)T->p2p[i].first;
    p2p[2 * i + 1] = (long long)T->p2p[i].second;
  }
}

// Field of order-m sources by the FMM on a prebuilt tree (tree.cc:359-385). F is overwritten.
int orc_fmm_field(void* t, int m, int n, const double* pos, const double* S, const double* alpha_iso,
                  const int* excl_off, const int* excl_idx, double damping, double* F) {
  Tree* T = static_cast<Tree*>(t);
  if (T->order - m < 0 || T->order > 12) {
    g_err = "unsupported expansion order";
    return 1;
  }
  Particles P;
  P.n = n;
  P.m = m;
  P.pos = pos;
  P.S = S;
  P.alpha = alpha_iso;
  P.excl_off = excl_off;
  P.excl_idx = excl_idx;
  fmm_field(*T, P, damping, F);
  return 0;
}

// Single FMM operators, exposed for operator-level tests (full multi-index arrays).
void orc_coulomb_derivs(int p, double x, double y, double z, double* D) {
  MITable T;
  T.init(p + 1);
  std::vector<double> v;
  coulomb_derivs(T, p, x, y, z, v);
  for (int i = 0; i < nterms(p); ++i) D[i] = v[i];
}

void orc_pair_field(int m, double dx, double dy, double dz, const double* S, double a, int damped,
                    double* F) {
  pair_field(m, dx, dy, dz, S, a, damped != 0, F);
}

void orc_thole_factors(double v, double* out) {
  Screen s = thole_factors(v);
  out[0] = s.fV;
  out[1] = s.fE;
  out[2] = s.fT;
  out[3] = s.fD;
}

// NuclearFields::compute (electric_fields.cc:58-75)
void orc_nuclear_fields(int n, const double* pos, int natoms, const double* charges,
                        const double* apos, double* F) {
#pragma omp parallel for
  for (int i = 0; i < n; ++i) {
    double f[3] = {0, 0, 0};
    for (int a = 0; a < natoms; ++a) {
      double d[3] = {pos[3 * i] - apos[3 * a], pos[3 * i + 1] - apos[3 * a + 1],
                     pos[3 * i + 2] - apos[3 * a + 2]};
      double R  = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      double r3 = 1.0 / (R * R * R);
      for (int c = 0; c < 3; ++c) f[c] -= charges[a] * (-d[c] * r3);  // T1 = -d/R^3
    }
    F[3 * i] = f[0];
    F[3 * i + 1] = f[1];
    F[3 * i + 2] = f[2];
  }
}

// MultipoleExpansion::interaction_energy (multipole_expansion.cc:10-33) with
// prefactors_nuclei (math.cc:84-97). kmax[i] = highest multipole order present at site i (-1 none);
// mul holds, per site, 1 + 3 + 6 packed values (zeros where absent).
double orc_nuclear_multipole_energy(int n, const double* pos, const int* kmax, const double* mul,
                                    int natoms, const double* charges, const double* apos) {
  MITable T;
  T.init(4);
  double total = 0.0;
#pragma omp parallel for reduction(+ : total)
  for (int i = 0; i < n; ++i) {
    std::vector<double> D;
    for (int a = 0; a < natoms; ++a) {
      double dx = apos[3 * a] - pos[3 * i];
      double dy = apos[3 * a + 1] - pos[3 * i + 1];
      double dz = apos[3 * a + 2] - pos[3 * i + 2];
      coulomb_derivs(T, 2, dx, dy, dz, D);
      for (int k = 0; k <= kmax[i] && k <= 2; ++k) {
        double taylor = (k % 2 == 0 ? 1.0 : -1.0);
        for (int t = 2; t <= k; ++t) taylor /= t;
        const double* mk = &mul[10 * i + nterms(k - 1)];
        double e = 0.0;
        for (int q = 0; q < ncomp(k); ++q) {
          int idx = nterms(k - 1) + q;
          // trinomial symmetry factor k!/(a! b! c!) (math.cc:55-73)
          double fk = 1.0;
          for (int t = 2; t <= k; ++t) fk *= t;
          double sym = fk * T.inv_fact[idx];
          e += taylor * sym * mk[q] * D[idx];
        }
        total += e * charges[a];
      }
    }
  }
  return total;
}

// ---- response operator and CG -----------------------------------------------------------
void* orc_bsys_create(int n, const double* pos, const double* alpha6, const int* excl_off,
                      const int* excl_idx, double damping, int fmm, int ncrit, int order,
                      double theta, int rebuild_per_apply) {
  BSystem* B = new BSystem();
  B->n = n;
  B->pos = pos;
  B->alpha6 = alpha6;
  B->excl_off = excl_off;
  B->excl_idx = excl_idx;
  B->damping = damping;
  B->fmm = fmm;
  B->ncrit = ncrit;
  B->order = order;
  B->theta = theta;
  B->rebuild_per_apply = rebuild_per_apply;
  bsys_init(*B);
  return B;
}

void orc_bsys_free(void* b) { delete static_cast<BSystem*>(b); }

void orc_bsys_apply(void* b, const double* x, double* out) { bsys_apply(*static_cast<BSystem*>(b), x, out); }

void orc_bsys_apply_alpha(void* b, const double* x, double* out) {
  apply_alpha(*static_cast<BSystem*>(b), x, out);
}

void orc_bsys_apply_alpha_inv(void* b, const double* x, double* out) {  // bmatrix.cc:53-62
  BSystem* B = static_cast<BSystem*>(b);
  for (int i = 0; i < B->n; ++i)
    for (int r = 0; r < 3; ++r) {
      const double* ai = &B->alpha_inv[9 * i + 3 * r];
      out[3 * i + r] = ai[0] * x[3 * i] + ai[1] * x[3 * i + 1] + ai[2] * x[3 * i + 2];
    }
}

// Dense B (bmatrix.cc:64-92): diag alpha^-1, off-diagonal -T2 (damped with v = damp/(a_i a_j)^(1/6)).
// excl values are compared with the site's potfile index (index[j]), as the dense builder does.
void orc_bsys_dense(void* b, const int* index, double* out) {
  BSystem* B  = static_cast<BSystem*>(b);
  const int N = 3 * B->n;
#pragma omp parallel for
  for (int i = 0; i < B->n; ++i) {
    for (int j = 0; j < B->n; ++j) {
      bool excl = false;
      for (int k = B->excl_off[i]; k < B->excl_off[i + 1]; ++k)
        if (B->excl_idx[k] == index[j]) excl = true;
      if (excl || i == j) continue;
      double d[3] = {B->pos[3 * j] - B->pos[3 * i], B->pos[3 * j + 1] - B->pos[3 * i + 1],
                     B->pos[3 * j + 2] - B->pos[3 * i + 2]};
      double R = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      Screen s;
      if (B->damping > 0.0) {
        double v = B->damping / std::pow(B->alpha_iso[i] * B->alpha_iso[j], 1.0 / 6.0);
        s        = thole_factors(v * R);
      }
      double ir3 = 1.0 / (R * R * R), ir5 = ir3 / (R * R);
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) {
          double t2 = 3.0 * s.fT * d[a] * d[c] * ir5 - (a == c ? s.fE * ir3 : 0.0);
          out[(size_t)(3 * i + a) * N + 3 * j + c] = -t2;
        }
    }
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) out[(size_t)(3 * i + a) * N + 3 * i + c] = B->alpha_inv[9 * i + 3 * a + c];
  }
}

// InducedMoments::compute (electric_fields.cc:191-239). Returns 0 on convergence, 1 otherwise.
// norms receives the residual norm of each iteration (length maxiter).
int orc_bsys_solve(void* b, const double* rhs, const double* guess, int make_guess, double thresh,
                   int maxiter, double* x_out, int* iters, double* norms) {
  BSystem* B  = static_cast<BSystem*>(b);
  const int N = 3 * B->n;
  std::vector<double> x(N), r(N), z(N), p(N), Ap(N), tmp(N);
  std::vector<double> rhsv(rhs, rhs + N);
  if (make_guess)
    apply_alpha(*B, rhs, x.data());
  else
    std::copy(guess, guess + N, x.begin());
  bsys_apply(*B, x.data(), tmp.data());
  for (int i = 0; i < N; ++i) r[i] = rhs[i] - tmp[i];
  apply_alpha(*B, r.data(), z.data());
  p = z;
  bool converged = false;
  int k = 0;
  for (k = 0; k < maxiter; ++k) {
    bsys_apply(*B, p.data(), Ap.data());
    double rz = dot(r, z);
    double alpha_k = rz / dot(p, Ap);
    for (int i = 0; i < N; ++i) x[i] += alpha_k * p[i];
    for (int i = 0; i < N; ++i) r[i] -= alpha_k * Ap[i];
    double rnorm = std::sqrt(dot(r, r));
    norms[k] = rnorm;
    if (rnorm < thresh) {
      converged = true;
      break;
    }
    apply_alpha(*B, r.data(), z.data());
    double beta_k = dot(z, r) / rz;
    for (int i = 0; i < N; ++i) p[i] = z[i] + beta_k * p[i];
  }
  std::copy(x.begin(), x.end(), x_out);
  *iters = converged ? k + 1 : maxiter;
  if (!converged) {
    g_err = "Failed to converge induced dipole moments.";
    return 1;
  }
  return 0;
}

}  // extern "C"
