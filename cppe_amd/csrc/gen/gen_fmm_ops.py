"""Generate straight-line Cartesian FMM operators for the gfx950 kernels (fmm_ops.gen.hh).

Mathematics (multi-index notation, d^k/k! = x^a y^b z^c / (a! b! c!)):
  multipole about z     M_n(z) = sum_b (z - x_b)^(n-j)/(n-j)! S_j(b)      (point source order |j|)
  M2M (child -> parent) M_n += sum_{k<=n} d^k/k! Mc_{n-k},        d = z_parent - z_child
  M2L (source -> target)L_n += sum_j M_j D_{n+j}(d),  |n|+|j| <= p,  d = z_target - z_source
  L2L (parent -> child) L_n += sum_k d^k/k! Lp_{n+k},              d = z_child - z_parent
  L2P                   F_a -= sum_k h^k/k! L_{k+e_a},             h = x - z_cell
The derivative tensors are harmonic (Laplace's equation for 1/|r|): D_{s+2e_z} = -D_{s+2e_x} - D_{s+2e_y}
for every s. Applied repeatedly to the source index, every D_{n+j} with j_z >= 2 becomes a fixed
integer combination of D_{n+j'} with j'_z < 2 and |j'| = |j|, whatever n. So before M2L each cell's
M is folded onto the 2l+1 components of every order l with j_z < 2:
  fold                  Mf_{j'} = sum_j c(j -> j') M_j,   L_n += sum_{j'} Mf_{j'} D_{n+j'}
which gives the same L with 322 instead of 406 multiply-adds per dipole M2L at p = 5. The local
coefficients are derivatives of a harmonic function as well, L_{n+2e_z} = -L_{n+2e_x} - L_{n+2e_y}
(exactly, term by term of the truncated sum), so M2L computes only the L_n with n_z < 2 and
  unfold                L_n = -L_{n-2e_z+2e_x} - L_{n-2e_z+2e_y}   (n_z >= 2, in increasing n_z)
restores the rest per cell before L2L: 265 multiply-adds per dipole M2L at p = 5.
with D_k(r) = d^k/dr^k (1/|r|) from the Hermite-Coulomb recursion of McMurchie & Davidson
(J. Comput. Phys. 26, 218 (1978)) in its point-charge limit:
  R^(j)_000 = (-1)^j (2j-1)!! / |r|^(2j+1),  R^(j)_{t+1,u,v} = t R^(j+1)_{t-1,u,v} + x R^(j+1)_{t,u,v}.

Two operator families ("kinds"):
  m1   dipole sources: M orders 1..p, L orders 0..p-1 (the induced-dipole field)
  mc   combined sources of orders 0..2 (charge, dipole, quadrupole): M and L orders 0..p
Multi-indices are ordered by total order, then x exponent descending, then y descending.

Run:  python cppe_amd/csrc/gen/gen_fmm_ops.py > cppe_amd/csrc/device/fmm_ops.gen.hh
"""
import sys
from math import factorial

ORDERS = range(2, 9)


def multi(k):
    return [(a, b, k - a - b) for a in range(k, -1, -1) for b in range(k - a, -1, -1)]


def table(p):
    out = []
    for k in range(p + 1):
        out += multi(k)
    return out


def nterms(p):
    return 0 if p < 0 else (p + 1) * (p + 2) * (p + 3) // 6


def lit(v):
    if float(v).is_integer():
        return "R(%d.0)" % int(v)
    return "R(%r)" % float(v)


class Kind:
    def __init__(self, name, msrc_min, p):
        self.name = name
        self.p = p
        self.mmin = msrc_min               # lowest multipole order kept in M
        self.mlist = [m for m in table(p) if sum(m) >= msrc_min]
        self.lord = p - msrc_min           # highest local order kept in L
        self.llist = table(self.lord)
        self.midx = {m: i for i, m in enumerate(self.mlist)}
        self.lidx = {m: i for i, m in enumerate(self.llist)}
        self.flist = [m for m in self.mlist if m[2] < 2]  # folded M components (stored first)
        self.fidx = {m: i for i, m in enumerate(self.flist)}
        self.lkeep = [n for n in self.llist if n[2] < 2]  # L components computed by M2L


def fold_coeffs(j):
    """{j': c} with D_{s+j} = sum c D_{s+j'} (j'_z < 2) for harmonic D."""
    if j[2] < 2:
        return {j: 1}
    out = {}
    for jj in ((j[0] + 2, j[1], j[2] - 2), (j[0], j[1] + 2, j[2] - 2)):
        for k, v in fold_coeffs(jj).items():
            out[k] = out.get(k, 0) - v
    return {k: v for k, v in out.items() if v}


def emit_monomials(var, pref, maxo, lines):
    """pref_x{i} = x^i/i! etc., and mono(a,b,c) products on demand."""
    for ax, v in zip("xyz", var):
        lines.append(f"  const R {pref}{ax}0 = R(1.0);")
        for i in range(1, maxo + 1):
            if i == 1:
                lines.append(f"  const R {pref}{ax}1 = {v};")
            else:
                lines.append(f"  const R {pref}{ax}{i} = {pref}{ax}{i-1} * {v} * {lit(1.0 / i)};")


def mono(pref, k):
    a, b, c = k
    terms = [f"{pref}x{a}" if a else None, f"{pref}y{b}" if b else None, f"{pref}z{c}" if c else None]
    terms = [t for t in terms if t]
    return " * ".join(terms) if terms else "R(1.0)"


def gen_p2m(K):
    """M_n += sum_{j src, j<=n} d^(n-j)/(n-j)! S_j ; sources: m1 -> S[0..2] (order 1);
    mc -> S[0] (q), S[1..3] (order-1), S[4..9] (order-2)."""
    src = multi(1) if K.name == "m1" else multi(0) + multi(1) + multi(2)
    lines = [f"template <typename R>\n__device__ __forceinline__ void p2m_{K.name}_p{K.p}(R dx, R dy, R dz, const R* __restrict__ S, R* __restrict__ M) {{"]
    emit_monomials(["dx", "dy", "dz"], "d", K.p, lines)
    for n in K.mlist:
        terms = []
        for q, j in enumerate(src):
            k = (n[0] - j[0], n[1] - j[1], n[2] - j[2])
            if min(k) < 0:
                continue
            terms.append(f"{mono('d', k)} * S[{q}]")
        if terms:
            lines.append(f"  M[{K.midx[n]}] += " + " + ".join(terms) + ";")
    lines.append("}")
    return "\n".join(lines)


def gen_m2m(K):
    lines = [f"template <typename R>\n__device__ __forceinline__ void m2m_{K.name}_p{K.p}(R dx, R dy, R dz, const R* __restrict__ Mc, R* __restrict__ Mp) {{"]
    emit_monomials(["dx", "dy", "dz"], "d", K.p, lines)
    for n in K.mlist:
        terms = []
        for ka in range(n[0] + 1):
            for kb in range(n[1] + 1):
                for kc in range(n[2] + 1):
                    j = (n[0] - ka, n[1] - kb, n[2] - kc)
                    if sum(j) < K.mmin:
                        continue
                    k = (ka, kb, kc)
                    m = mono("d", k)
                    terms.append(f"Mc[{K.midx[j]}]" if m == "R(1.0)" else f"{m} * Mc[{K.midx[j]}]")
        lines.append(f"  Mp[{K.midx[n]}] += " + " + ".join(terms) + ";")
    lines.append("}")
    return "\n".join(lines)


def gen_derivs(p, lines, rinv_expr):
    """D_k for |k| <= p as constants D_a_b_c via the Hermite-Coulomb recursion."""
    lines.append(f"  const R rinv = {rinv_expr};")
    lines.append("  const R rinv2 = rinv * rinv;")
    # base_j = (-1)^j (2j-1)!! rinv^(2j+1)
    lines.append("  const R h0_0_0_0 = rinv;")
    for j in range(1, p + 1):
        lines.append(f"  const R h{j}_0_0_0 = h{j-1}_0_0_0 * rinv2 * {lit(-(2 * j - 1))};")
    for k in range(1, p + 1):
        for j in range(0, p - k + 1):
            for (t, u, v) in multi(k):
                if t > 0:
                    e = f"dx * h{j+1}_{t-1}_{u}_{v}"
                    if t - 1 > 0:
                        e += f" + {lit(t - 1)} * h{j+1}_{t-2}_{u}_{v}"
                elif u > 0:
                    e = f"dy * h{j+1}_{t}_{u-1}_{v}"
                    if u - 1 > 0:
                        e += f" + {lit(u - 1)} * h{j+1}_{t}_{u-2}_{v}"
                else:
                    e = f"dz * h{j+1}_{t}_{u}_{v-1}"
                    if v - 1 > 0:
                        e += f" + {lit(v - 1)} * h{j+1}_{t}_{u}_{v-2}"
                lines.append(f"  const R h{j}_{t}_{u}_{v} = {e};")


def gen_fold(K):
    """In place: M[0 .. |flist|) <- folded components (the rest of the row is left as is)."""
    lines = [f"template <typename R>\n__device__ __forceinline__ void fold_{K.name}_p{K.p}(R* __restrict__ M) {{"]
    acc = {jp: [] for jp in K.flist}
    for j in K.mlist:
        for jp, c in fold_coeffs(j).items():
            acc[jp].append((c, j))
    used = sorted({K.midx[j] for jp in K.flist for _, j in acc[jp]})
    for i in used:
        lines.append(f"  const R m{i} = M[{i}];")
    for jp in K.flist:
        (c0, j0), rest = acc[jp][0], acc[jp][1:]
        assert c0 == 1 and j0 == jp
        e = f"m{K.midx[jp]}"
        for c, j in rest:
            v = f"m{K.midx[j]}"
            mag = v if abs(c) == 1 else f"{lit(abs(c))} * {v}"
            e += (" + " if c > 0 else " - ") + mag
        lines.append(f"  M[{K.fidx[jp]}] = {e};")
    lines.append("}")
    return "\n".join(lines)


def gen_m2l(K, prec):
    rinv = "rsqrt_f64(dx * dx + dy * dy + dz * dz)" if prec == "f64" else "rsqrt_f32(dx * dx + dy * dy + dz * dz)"
    R = "double" if prec == "f64" else "float"
    lines = [f"__device__ __forceinline__ void m2l_{K.name}_p{K.p}_{prec}({R} dx, {R} dy, {R} dz, const {R}* __restrict__ M, {R}* __restrict__ L) {{",
             f"  using R = {R};"]
    gen_derivs(K.p, lines, rinv)
    # contraction over the folded source (gen_fold), ordered by source coefficient: each Mf_j is
    # loaded once and scattered into every L_n it reaches, so only the derivatives and the L
    # accumulators stay live
    for j in K.flist:
        terms = []
        for n in K.lkeep:
            s = (n[0] + j[0], n[1] + j[1], n[2] + j[2])
            if sum(s) <= K.p:
                terms.append(f"L[{K.lidx[n]}] += m * h0_{s[0]}_{s[1]}_{s[2]};")
        if terms:
            lines.append(f"  {{ const R m = M[{K.fidx[j]}]; " + " ".join(terms) + " }")
    lines.append("}")
    return "\n".join(lines)


def gen_unfold(K):
    """In place: the L_n with n_z >= 2 from those with smaller n_z (harmonic local expansion)."""
    lines = [f"template <typename R>\n__device__ __forceinline__ void unfold_{K.name}_p{K.p}(R* __restrict__ L) {{"]
    for n in sorted((n for n in K.llist if n[2] >= 2), key=lambda n: n[2]):
        a = K.lidx[(n[0] + 2, n[1], n[2] - 2)]
        b = K.lidx[(n[0], n[1] + 2, n[2] - 2)]
        lines.append(f"  L[{K.lidx[n]}] = -(L[{a}] + L[{b}]);")
    lines.append("}")
    kept = ", ".join("true" if n[2] < 2 else "false" for n in K.llist)
    lines.append(f"constexpr bool lkept_{K.name}_p{K.p}(int k) {{")
    lines.append(f"  constexpr bool t[] = {{{kept}}};")
    lines.append("  return t[k];")
    lines.append("}")
    return "\n".join(lines)


def gen_l2l(K):
    lines = [f"template <typename R>\n__device__ __forceinline__ void l2l_{K.name}_p{K.p}(R dx, R dy, R dz, const R* __restrict__ Lp, R* __restrict__ Lc) {{"]
    emit_monomials(["dx", "dy", "dz"], "d", K.lord, lines)
    for n in K.llist:
        terms = []
        for k in table(K.lord - sum(n)):
            s = (n[0] + k[0], n[1] + k[1], n[2] + k[2])
            m = mono("d", k)
            terms.append(f"Lp[{K.lidx[s]}]" if m == "R(1.0)" else f"{m} * Lp[{K.lidx[s]}]")
        lines.append(f"  Lc[{K.lidx[n]}] += " + " + ".join(terms) + ";")
    lines.append("}")
    return "\n".join(lines)


def gen_l2p(K):
    lines = [f"template <typename R>\n__device__ __forceinline__ void l2p_{K.name}_p{K.p}(R hx, R hy, R hz, const R* __restrict__ L, R& fx, R& fy, R& fz) {{"]
    emit_monomials(["hx", "hy", "hz"], "h", max(K.lord - 1, 0), lines)
    for a, f in enumerate(["fx", "fy", "fz"]):
        terms = []
        for k in table(K.lord - 1):
            s = list(k)
            s[a] += 1
            m = mono("h", k)
            idx = K.lidx[tuple(s)]
            terms.append(f"L[{idx}]" if m == "R(1.0)" else f"{m} * L[{idx}]")
        if terms:
            lines.append(f"  {f} -= " + " + ".join(terms) + ";")
    lines.append("}")
    return "\n".join(lines)


def main():
    out = ["// GENERATED by cppe_amd/csrc/gen/gen_fmm_ops.py -- do not edit by hand.",
           "// Straight-line Cartesian FMM operators (see the generator's docstring for the mathematics).",
           "#pragma once", "#ifndef CPPE_FMMOPS_HOST", "#include \"common.hh\"", "#endif", "",
           "namespace cppe_dev {", "namespace fmmops {", ""]
    sizes = []
    for p in ORDERS:
        for name, mmin in (("m1", 1), ("mc", 0)):
            K = Kind(name, mmin, p)
            sizes.append((name, p, len(K.mlist), len(K.llist)))
            out.append("#if !defined(CPPE_FMMOPS_ONLY_P) || CPPE_FMMOPS_ONLY_P == %d" % p)
            out.append(gen_p2m(K))
            out.append(gen_m2m(K))
            out.append(gen_fold(K))
            out.append(gen_m2l(K, "f64"))
            out.append(gen_m2l(K, "f32"))
            out.append(gen_unfold(K))
            out.append(gen_l2l(K))
            out.append(gen_l2p(K))
            out.append("#endif")
            out.append("")
    out.append("}  // namespace fmmops")
    out.append("")
    out.append("// (kind, order) -> number of M and L coefficients")
    out.append("constexpr int fmm_msize(int kind_mc, int p) {")
    out.append("  return kind_mc ? (p + 1) * (p + 2) * (p + 3) / 6 : (p + 1) * (p + 2) * (p + 3) / 6 - 1;")
    out.append("}")
    out.append("// folded M components (orders kept in M, 2l + 1 per order l)")
    out.append("constexpr int fmm_fsize(int kind_mc, int p) {")
    out.append("  return kind_mc ? (p + 1) * (p + 1) : (p + 1) * (p + 1) - 1;")
    out.append("}")
    out.append("constexpr int fmm_lsize(int kind_mc, int p) {")
    out.append("  return kind_mc ? (p + 1) * (p + 2) * (p + 3) / 6 : p * (p + 1) * (p + 2) / 6;")
    out.append("}")
    out.append("constexpr int kFmmMinOrder = %d, kFmmMaxOrder = %d;" % (min(ORDERS), max(ORDERS)))
    out.append("}  // namespace cppe_dev")
    sys.stdout.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
