"""Generator for the pointwise BSSN right-hand side used by the HIP kernels.

Restates the equations of the reference generator (CodeGen/bssn.py:78-218,
standard gauge, constant eta damping, unstaged) with the tensor helpers of
CodeGen/dendro.py:173-550. Unlike the reference, which expands every equation
fully before eliminating common subexpressions, the equations are built
through named tensor intermediates (inverse metric, Christoffel symbols,
Ricci tensor, raised At, ...), which stay atoms; common subexpressions are then
eliminated around them and all statements are ordered by a register-pressure
list scheduler. The result keeps the live set small enough for FP64 on a
64-wide wavefront.

The emitted function reads inputs through an accessor object
(`in.template fld<V>()`, `in.template d1<V, D>()`, `in.template d2<M, C>()`)
and hands each right-hand side to `out.template put<V>(value)`, so a kernel
decides where inputs come from and what happens to each output.

Usage: python gen_bssn_rhs.py [--cse=local|global] [--phased | --ham | --cons-phased [--search]] [--gauge]
       [output.h]
(--phased: the right-hand side in the phase order of the fused kernel;
 --gauge: with or without --phased, the right-hand side with the gauge variants
 of the reference's other builds, the Rochester gauge and the slow-start lapse;
 --ham: the constraints (Hamiltonian, momentum, psi4);
 --cons-phased: the constraints in the phase order of the fused constraint
 kernel, --search to look for that order again)
(local: common subexpressions are shared within one equation or intermediate
only, which bounds the live set; global: shared across all of them.)
"""
from __future__ import annotations

import sys
import time

import sympy as sp
from sympy import Rational
from sympy.printing.c import C99CodePrinter
from sympy.printing.precedence import PRECEDENCE

# ---------------------------------------------------------------------------
# variables (BSSN_GR/include/grDef.h:35-60 order)
VARS = ["alpha", "chi", "K", "Gt0", "Gt1", "Gt2", "beta0", "beta1", "beta2", "B0", "B1", "B2",
        "gt0", "gt1", "gt2", "gt3", "gt4", "gt5", "At0", "At1", "At2", "At3", "At4", "At5"]
VIDX = {n: i for i, n in enumerate(VARS)}
D2VARS = ["alpha", "chi", "beta0", "beta1", "beta2", "gt0", "gt1", "gt2", "gt3", "gt4", "gt5"]
D2IDX = {n: i for i, n in enumerate(D2VARS)}
SYM6 = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
PAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
E_I = [0, 1, 2]
E_IJ = [(i, j) for i in E_I for j in E_I]

fld = {n: sp.Symbol(f"f_{n}", real=True) for n in VARS}
d1s = {(n, d): sp.Symbol(f"d1_{n}_{d}", real=True) for n in VARS for d in E_I}
d2s = {(n, c): sp.Symbol(f"d2_{n}_{c}", real=True) for n in D2VARS for c in range(6)}
XYZ = [sp.Symbol(f"x_{i}", real=True) for i in range(3)]  # point coordinates (psi4 only)
INPUTS = set(fld.values()) | set(d1s.values()) | set(d2s.values()) | set(XYZ)
eta = sp.Symbol("eta", real=True)
l1, l2, l3, l4 = sp.symbols("lam0 lam1 lam2 lam3", real=True)
lf0, lf1 = sp.symbols("lamf0 lamf1", real=True)

DEFS: list[tuple[sp.Symbol, sp.Expr]] = []  # named intermediates, in definition order


def let(name, expr):
    s = sp.Symbol(name, real=True)
    DEFS.append((s, sp.sympify(expr)))
    return s


def d(i, name):
    return d1s[(name, i)]


def dd(i, j, name):
    return d2s[(name, SYM6[min(i, j)][max(i, j)])]


def sym_tensor(prefix, fn):
    """Symmetric 3x3 of named intermediates: fn(i, j) for i <= j."""
    s = {}
    for i, j in PAIRS:
        s[(i, j)] = s[(j, i)] = let(f"{prefix}{i}{j}", fn(i, j))
    return sp.Matrix(3, 3, lambda i, j: s[(i, j)])


a, chi, K = fld["alpha"], fld["chi"], fld["K"]
Gt_n = [f"Gt{i}" for i in E_I]
b_n = [f"beta{i}" for i in E_I]
B_n = [f"B{i}" for i in E_I]
gt_n = [[f"gt{SYM6[i][j]}" for j in E_I] for i in E_I]
At_n = [[f"At{SYM6[i][j]}" for j in E_I] for i in E_I]
b = [fld[n] for n in b_n]
B = [fld[n] for n in B_n]
gt = sp.Matrix(3, 3, lambda i, j: fld[gt_n[i][j]])
At = sp.Matrix(3, 3, lambda i, j: fld[At_n[i][j]])

# inverse conformal metric: adjugate over determinant (dendro.py:375-393)
cof = {(0, 0): gt[1, 1] * gt[2, 2] - gt[1, 2] ** 2, (0, 1): gt[0, 2] * gt[1, 2] - gt[0, 1] * gt[2, 2],
       (0, 2): gt[0, 1] * gt[1, 2] - gt[0, 2] * gt[1, 1], (1, 1): gt[0, 0] * gt[2, 2] - gt[0, 2] ** 2,
       (1, 2): gt[0, 1] * gt[0, 2] - gt[0, 0] * gt[1, 2], (2, 2): gt[0, 0] * gt[1, 1] - gt[0, 1] ** 2}
idet = let("idet", 1 / (gt[0, 0] * cof[(0, 0)] + gt[0, 1] * cof[(0, 1)] + gt[0, 2] * cof[(0, 2)]))
igt = sym_tensor("igt", lambda i, j: cof[(i, j)] * idet)
ichi = let("ichi", 1 / chi)

# first Christoffel C1_kij = 1/2 (d_j g_ki + d_i g_kj - d_k g_ij) (dendro.py:418-421)
C1 = {}
for k in E_I:
    for i, j in PAIRS:
        C1[(k, i, j)] = C1[(k, j, i)] = let(
            f"C1_{k}{i}{j}", sp.Float(0.5) * (d(j, gt_n[k][i]) + d(i, gt_n[k][j]) - d(k, gt_n[i][j])))
# second Christoffel C2^i_jk = igt^il C1_ljk (dendro.py:445-446)
C2 = {}
for i in E_I:
    for j, k in PAIRS:
        C2[(i, j, k)] = C2[(i, k, j)] = let(f"C2_{i}{j}{k}", sum(igt[i, l] * C1[(l, j, k)] for l in E_I))
# contracted connection (dendro.py:512)
CalGt = [let(f"CalGt{i}", sum(igt[k, l] * C2[(i, k, l)] for k, l in E_IJ)) for i in E_I]


def C3(i, j, k):  # physical-metric Christoffel (dendro.py:471-480)
    return C2[(i, j, k)] - sp.Float(0.5) * ichi * (
        (d(k, "chi") if i == j else 0) + (d(j, "chi") if i == k else 0)
        - gt[j, k] * sum(igt[i, m] * d(m, "chi") for m in E_I))


# D_i D_j alpha (dendro.py:181-183)
DDa = sym_tensor("DDa", lambda i, j: dd(i, j, "alpha") - sum(C3(l, i, j) * d(l, "alpha") for l in E_I))

# Ricci tensor R = Rt + Rphi (dendro.py:495-541); 1/chi written as ichi
lap_chi = let("lapchi", sum(igt[k, l] * (dd(k, l, "chi") - Rational(3, 2) * ichi * d(k, "chi") * d(l, "chi"))
                            for k, l in E_IJ) - sum(CalGt[m] * d(m, "chi") for m in E_I))


def ricci(i, j):
    rt = (-sp.Float(0.5) * sum(igt[l, m] * dd(l, m, gt_n[i][j]) for l, m in E_IJ)
          + sp.Float(0.5) * sum(gt[k, i] * d(j, Gt_n[k]) + gt[k, j] * d(i, Gt_n[k]) for k in E_I)
          + sp.Float(0.5) * sum(CalGt[k] * (C1[(i, j, k)] + C1[(j, i, k)]) for k in E_I)
          + sum(igt[l, m] * (C2[(k, l, i)] * C1[(j, k, m)] + C2[(k, l, j)] * C1[(i, k, m)]
                             + C2[(k, i, m)] * C1[(k, l, j)]) for k in E_I for l, m in E_IJ))
    xrphi = (Rational(1, 2) * ichi * (dd(i, j, "chi") - sum(C2[(k, j, i)] * d(k, "chi") for k in E_I))
             - Rational(1, 4) * ichi * ichi * d(i, "chi") * d(j, "chi"))
    return rt + xrphi + Rational(1, 2) * ichi * gt[i, j] * lap_chi


R = sym_tensor("R", ricci)
# raised At (dendro.py:203-208)
AtUU = sym_tensor("AtUU", lambda i, j: sum(igt[i, k] * igt[j, l] * At[k, l] for k, l in E_IJ))
divb = let("divb", sum(d(j, b_n[j]) for j in E_I))


def lie_scalar(f_name, weight=0):  # dendro.py:230-231
    return sum(b[i] * d(i, f_name) for i in E_I) + weight * fld[f_name] * divb


def lie_sym(names, M, weight):  # dendro.py:236-239
    return sp.Matrix(3, 3, lambda i, j: sum(
        b[k] * d(k, names[i][j]) + M[i, k] * d(j, b_n[k]) + M[k, j] * d(i, b_n[k]) for k in E_I)
        + weight * M[i, j] * divb)


weight = -Rational(2, 3)  # bssn.py:39

a_rhs = l1 * lie_scalar("alpha") - 2 * a * K
b_rhs = [Rational(3, 4) * (lf0 + lf1 * a) * B[i] + l2 * sum(b[j] * d(j, b_n[i]) for j in E_I)
         for i in E_I]
gt_rhs = lie_sym(gt_n, gt, weight) - 2 * a * At
chi_rhs = lie_scalar("chi", weight) + Rational(2, 3) * (chi * a * K)
X = sp.Matrix(3, 3, lambda i, j: a * R[i, j] - DDa[i, j])
trX = let("trX", sum(igt[i, j] * X[i, j] for i, j in E_IJ))  # trace_free, dendro.py:352-365
AikAkj = sp.Matrix(3, 3, lambda i, j: sum(At[i, k] * sum(igt[k, l] * At[l, j] for l in E_I)
                                          for k in E_I))
At_rhs = lie_sym(At_n, At, weight) + chi * sp.Matrix(3, 3, lambda i, j: X[i, j] - gt[i, j] * trX / 3) \
    + a * (K * At - 2 * AikAkj)
# K: laplacian of alpha with the physical inverse metric chi * igt (dendro.py:288-290)
K_rhs = (sum(b[i] * d(i, "K") for i in E_I) - chi * sum(igt[i, j] * DDa[i, j] for i, j in E_IJ)
         + a * (K * K / 3 + sum(At[i, j] * AtUU[i, j] for i, j in E_IJ)))
Gt_rhs = [
    sum(b[j] * d(j, Gt_n[i]) for j in E_I)
    - sum(CalGt[j] * d(j, b_n[i]) for j in E_I)
    + Rational(2, 3) * CalGt[i] * divb
    + sum(igt[j, k] * dd(j, k, b_n[i]) + igt[i, j] * dd(j, k, b_n[k]) / 3 for j, k in E_IJ)
    - sum(2 * AtUU[i, j] * d(j, "alpha") for j in E_I)
    + sum(2 * a * C2[(i, j, k)] * AtUU[j, k] for j, k in E_IJ)
    - sum(a * (3 * ichi * AtUU[i, j] * d(j, "chi") + Rational(4, 3) * igt[i, j] * d(j, "K"))
          for j in E_I)
    for i in E_I]
Gt_sym = [let(f"Gtrhs{i}", Gt_rhs[i]) for i in E_I]
B_rhs = [Gt_sym[i] - eta * B[i] + l3 * sum(b[j] * d(j, B_n[i]) for j in E_I)
         - l4 * sum(b[j] * d(j, Gt_n[i]) for j in E_I) for i in E_I]

OUTPUTS = [None] * 24
OUTPUTS[VIDX["alpha"]] = a_rhs
OUTPUTS[VIDX["chi"]] = chi_rhs
OUTPUTS[VIDX["K"]] = K_rhs
for i in E_I:
    OUTPUTS[VIDX[f"Gt{i}"]] = Gt_sym[i]
    OUTPUTS[VIDX[f"beta{i}"]] = b_rhs[i]
    OUTPUTS[VIDX[f"B{i}"]] = B_rhs[i]
for i, j in PAIRS:
    OUTPUTS[VIDX["gt0"] + SYM6[i][j]] = gt_rhs[i, j]
    OUTPUTS[VIDX["At0"] + SYM6[i][j]] = At_rhs[i, j]

# The gauge variants of the reference's other generated builds (rhs.cpp:219-265),
# chosen at run time:
# * the Rochester puncture gauge (bssn.py:415-455, bssn_rochester_puncture_gauge):
#   b_rhs^i = xi_2 beta^j d_j beta^i + 3/4 xi_3 Gt^i - eta beta^i and B frozen
#   (B_rhs = 0; the generated bssneqs_*_rochester_gauge*.cpp write no B_rhs at
#   all); xi_2, xi_3 are BSSN_XI[1], BSSN_XI[2]. SEL(flag, x, y) prints as a
#   select, so the standard gauge's values do not see the Rochester terms (an
#   infinite eta times 0 would otherwise turn them into NaN);
# * the slow-start lapse (bssn.py:90-99, sslGaugeCondition): a_rhs gains
#   -sslt with sslt = 0.6 sqrt(chi) (alpha - sqrt(chi)) exp(-t^2/800), which
#   the caller forms (0 without the term).
xi1, xi2 = sp.symbols("xi1 xi2", real=True)
rochester = sp.Symbol("rochester", real=True)
sslt = sp.Symbol("sslt", real=True)
SEL = sp.Function("SEL")
OUTPUTS_GAUGE = list(OUTPUTS)
OUTPUTS_GAUGE[VIDX["alpha"]] = a_rhs - sslt
for i in E_I:
    b_roch = (xi1 * sum(b[j] * d(j, b_n[i]) for j in E_I) + Rational(3, 4) * xi2 * fld[Gt_n[i]]
              - eta * b[i])
    OUTPUTS_GAUGE[VIDX[f"beta{i}"]] = SEL(rochester, b_roch, b_rhs[i])
    OUTPUTS_GAUGE[VIDX[f"B{i}"]] = SEL(rochester, sp.Integer(0), B_rhs[i])


# Hamiltonian constraint (CodeGen/constraints.py:250-254; dendro.sqr, dendro.py:309-318):
# chi igt^jk R_jk - At_ij At^ij + 2/3 K^2, with the same R as the right-hand side.
HAM = (sum(chi * igt[j, k] * R[j, k] for j, k in E_IJ) - sum(At[i, j] * AtUU[i, j] for i, j in E_IJ)
       + Rational(2, 3) * K ** 2)

# Momentum constraints (CodeGen/constraints.py:256-282), with the evolved Gt^j:
# igt^jk (d_k At_ij - C2^m_ki At_jm) - Gt^j At_ij - 3/2 igt^jk At_ki d_j chi / chi
# - 2/3 d_i K
MOM = [sum(igt[j, k] * (d(k, At_n[i][j]) - sum(C2[(m, k, i)] * At[j, m] for m in E_I)) for j, k in E_IJ)
       - sum(fld[Gt_n[j]] * At[i, j] for j in E_I)
       - Rational(3, 2) * sum(igt[j, k] * At[k, i] * d(j, "chi") for j, k in E_IJ) * ichi
       - Rational(2, 3) * d(i, "K") for i in E_I]


def psi4_exprs():
    """psi4 (CodeGen/constraints.py:42-244): Gram-Schmidt in the physical
    metric gt/chi on r = (x,y,z), theta = (xz, yz, -(x^2+y^2)) and
    phi = (-y, x, 0) gives the tetrad r, m = (theta + i phi)/sqrt(2). Every
    term is quadratic in m, so the two factors 1/sqrt(2) are applied as one
    1/2. Returns [real, imaginary] with the sign convention of
    constraints.py:238-244. The intermediates are created here (not at import)
    so that the right-hand side's definition list stays without them."""
    x, y, z = XYZ
    half = Rational(1, 2)

    def gdot(u, v):
        return ichi * sum(gt[i, j] * u[i] * v[j] for i, j in E_IJ)

    r0 = [x, y, z]
    nr = let("ps_nr", 1 / sp.sqrt(gdot(r0, r0)))
    r = [let(f"ps_r{i}", r0[i] * nr) for i in E_I]
    th0 = [x * z, y * z, -(x * x + y * y)]
    t1, t2 = let("ps_t1", gdot(th0, th0)), let("ps_t2", gdot(r, th0))
    nth = let("ps_nth", 1 / sp.sqrt(t1 - t2 * t2))
    th = [let(f"ps_th{i}", (th0[i] - t2 * r[i]) * nth) for i in E_I]
    ph0 = [-y, x, sp.Integer(0)]
    p1, p2, p3 = let("ps_p1", gdot(ph0, ph0)), let("ps_p2", gdot(r, ph0)), let("ps_p3", gdot(th, ph0))
    nph = let("ps_nph", 1 / sp.sqrt(p1 - p2 * p2 - p3 * p3))
    ph = [let(f"ps_ph{i}", (ph0[i] - p2 * r[i] - p3 * th[i]) * nph) for i in E_I]
    # 2 MM_ij, 2 NN_ij (symmetric); sqrt(2) MR_ij, sqrt(2) NR_ij (antisymmetric)
    MM = {(i, j): th[i] * th[j] - ph[i] * ph[j] for i, j in E_IJ}
    NN = {(i, j): th[i] * ph[j] + th[j] * ph[i] for i, j in E_IJ}
    MR = {(i, j): th[i] * r[j] - th[j] * r[i] for i, j in E_IJ}
    NR = {(i, j): ph[i] * r[j] - ph[j] * r[i] for i, j in E_IJ}
    Av = [let(f"ps_A{i}", sum(At[i, j] * r[j] for j in E_I)) for i in E_I]
    # d_j At_ki + C2^m_ki At_mj, contracted with sqrt(2) m_real (U) and sqrt(2) m_imag (V)
    Wt = {(k, i, j): let(f"ps_W{k}{i}{j}", d(j, At_n[k][i]) + sum(C2[(m, k, i)] * At[m, j] for m in E_I))
          for k in E_I for i, j in E_IJ}
    U = {(i, j): sum(th[k] * Wt[(k, i, j)] for k in E_I) for i, j in E_IJ}
    V = {(i, j): sum(ph[k] * Wt[(k, i, j)] for k in E_I) for i, j in E_IJ}
    rdchi = sum(r[i] * d(i, "chi") for i in E_I)
    Atmp = let("ps_Atmp", ichi * ichi * (sum(Av[i] * r[i] for i in E_I) + K * chi * Rational(1, 3)
                                         + half * rdchi))
    p1r = half * sum(R[i, j] * MM[(i, j)] for i, j in E_IJ)
    p1i = half * sum(R[i, j] * NN[(i, j)] for i, j in E_IJ)
    p2r = Atmp * half * sum(At[i, j] * MM[(i, j)] for i, j in E_IJ)
    p2i = Atmp * half * sum(At[i, j] * NN[(i, j)] for i, j in E_IJ)
    p3r = half * ichi * sum(MR[(i, j)] * U[(i, j)] - NR[(i, j)] * V[(i, j)] for i, j in E_IJ)
    p3i = half * ichi * sum(NR[(i, j)] * U[(i, j)] + MR[(i, j)] * V[(i, j)] for i, j in E_IJ)
    tA = let("ps_tA", sum(th[i] * Av[i] for i in E_I))
    pA = let("ps_pA", sum(ph[i] * Av[i] for i in E_I))
    td = sum(th[i] * d(i, "chi") for i in E_I)
    pd = sum(ph[i] * d(i, "chi") for i in E_I)
    p4r = half * ichi * ichi * (tA * (tA + half * td) - pA * (pA + half * pd))
    p4i = half * ichi * ichi * (tA * (pA + half * pd) + pA * (tA + half * td))
    return [-(p1r + p2r - p3r - p4r), p1i + p2i - p3i - p4i]


# ---------------------------------------------------------------------------
# code emission

def input_expr(sym: sp.Symbol) -> str:
    n = sym.name
    if n.startswith("f_"):
        return f"in.template fld<{VIDX[n[2:]]}>()"
    if n.startswith("d1_"):
        var, dd_ = n[3:].rsplit("_", 1)
        return f"in.template d1<{VIDX[var]}, {dd_}>()"
    if n.startswith("d2_"):
        var, c = n[3:].rsplit("_", 1)
        return f"in.template d2<{D2IDX[var]}, {c}>()"
    if n.startswith("x_"):
        return f"in.template crd<{n[2:]}>()"
    raise KeyError(n)


PARAMS = {"eta": "eta", "lam0": "lam[0]", "lam1": "lam[1]", "lam2": "lam[2]", "lam3": "lam[3]",
          "lamf0": "lamf[0]", "lamf1": "lamf[1]", "xi1": "xi[1]", "xi2": "xi[2]", "sslt": "sslt"}


class Printer(C99CodePrinter):
    def __init__(self, names):
        super().__init__({"strict": False})
        self.names = names

    def _print_Symbol(self, s):
        if s in self.names:
            return self.names[s]
        if s.name in PARAMS:
            return PARAMS[s.name]
        raise KeyError(s.name)

    def _print_Function(self, e):
        if e.func.__name__ != "SEL":
            return super()._print_Function(e)
        flag, x, y = e.args
        assert flag == rochester
        return f"(rochester ? {self._print_sel_arm(x)} : {self._print_sel_arm(y)})"

    def _print_sel_arm(self, x):
        return "T(0.0)" if x == 0 else f"({self._print(x)})"

    def _print_Rational(self, r):
        return f"T({r.p}.0/{r.q}.0)" if r.q != 1 else f"T({r.p}.0)"

    def _print_Integer(self, i):
        return f"T({int(i)}.0)"

    def _print_Float(self, f):
        return f"T({float(f)!r})"

    def _print_Pow(self, e):
        base, ex = e.as_base_exp()
        bs = self.parenthesize(base, PRECEDENCE["Pow"])
        if ex == -1:
            return f"(T(1.0)/{bs})"
        if ex == Rational(1, 2):
            return f"sqrt({self._print(base)})"
        if ex == Rational(-1, 2):
            return f"(T(1.0)/sqrt({self._print(base)}))"
        if ex.is_Integer and 2 <= int(ex) <= 4:
            return "(" + "*".join([bs] * int(ex)) + ")"
        if ex.is_Integer and -4 <= int(ex) <= -2:
            return "(T(1.0)/(" + "*".join([bs] * (-int(ex))) + "))"
        raise ValueError(f"unsupported power {e}")


def schedule(defs, outs):
    """Register-pressure-aware order of all statements (input loads,
    temporaries, output stores). Bottom-up list scheduling: starting from the
    stores, repeatedly place (earliest-so-far) the ready statement that adds
    the fewest values to the live set; a statement becomes ready once every
    statement using its value is placed. Returns program order as a list of
    ('in', sym) / ('tmp', sym) / ('out', v)."""
    ops = {}

    def operands(e):
        return [("tmp", x) if x in defs else ("in", x)
                for x in e.free_symbols if x in defs or x in INPUTS]

    for s_, e in defs.items():
        ops[("tmp", s_)] = operands(e)
    for v, e in enumerate(outs):
        ops[("out", v)] = operands(e)
    for x in INPUTS:
        ops[("in", x)] = []
    users = {n: 0 for n in ops}
    for n, o in ops.items():
        for x in set(o):
            users[x] += 1
    ready = {n for n in ops if users[n] == 0 and n[0] != "in"}

    def key(n):
        return (n[0], n[1] if isinstance(n[1], int) else n[1].name)

    live = set()
    rev = []
    while ready:
        best, best_score = None, None
        for n in ready:
            new = sum(1 for x in set(ops[n]) if x not in live)
            score = (new - (1 if n in live else 0), -len(ops[n]), key(n))
            if best_score is None or score < best_score:
                best, best_score = n, score
        ready.discard(best)
        rev.append(best)
        live.discard(best)
        for x in set(ops[best]):
            live.add(x)
            users[x] -= 1
            if users[x] == 0:
                ready.add(x)
    return rev[::-1]


def _reachable_defs(outputs):
    """The named intermediates the given outputs depend on, in definition order."""
    by_sym = dict(DEFS)
    need, stack = set(), [x for e in outputs for x in sp.sympify(e).free_symbols]
    while stack:
        x = stack.pop()
        if x in by_sym and x not in need:
            need.add(x)
            stack.extend(by_sym[x].free_symbols)
    return [(s_, e) for s_, e in DEFS if s_ in need]


def eliminate(outputs=None):
    """Common-subexpression elimination; returns (definitions, outputs, op count)."""
    outs = [sp.sympify(e) for e in (OUTPUTS if outputs is None else outputs)]
    defs_in = DEFS if outputs is None else _reachable_defs(outs)
    def_syms = [s for s, _ in defs_in]
    exprs = [e for _, e in defs_in] + outs
    if CSE_MODE == "global":
        repl, reduced = sp.cse(exprs, symbols=sp.numbered_symbols("c"), optimizations="basic")
    else:  # per expression: sharing across expressions only through named intermediates
        ctr = sp.numbered_symbols("c")
        repl, reduced = [], []
        for e in exprs:
            r, rd = sp.cse([e], symbols=ctr, optimizations="basic")
            repl += r
            reduced.append(rd[0])
    defs = dict(repl)
    for s, e in zip(def_syms, reduced[:len(defs_in)]):
        defs[s] = e
    outs_r = reduced[len(defs_in):]
    ops = sum(sp.count_ops(e) for e in defs.values()) + sum(sp.count_ops(e) for e in outs_r)
    return defs, outs_r, ops


HEADER_TOP = """// Generated by dendro-gr_amd/codegen/gen_bssn_rhs.py -- do not edit.
//
// Pointwise BSSN right-hand side (standard gauge, constant eta damping): the
// equations of CodeGen/bssn.py:78-218 with the tensor helpers of
// CodeGen/dendro.py:173-550, built through named tensor intermediates with
// common subexpressions eliminated around them.
"""


def emit(order, defs, outs_r, input_fn, phase_line=None, inline_fields=False, pin=False, out_names=None):
    """C++ statements for a program order of ('in'|'tmp'|'out', x) nodes, with
    optional ('phase', p) markers; returns (lines, temporaries, max live).
    With inline_fields, every use of a field reads it again (the kernel keeps
    fields in memory rather than in registers across phases)."""
    names, lines = {}, []
    if inline_fields:
        for x in fld.values():
            names[x] = input_fn(x)
    pr = Printer(names)
    ntmp = 0
    last_use = {}
    for k, (kind, x) in enumerate(order):
        e = defs[x] if kind == "tmp" else outs_r[x] if kind == "out" else None
        if e is not None:
            for y in e.free_symbols:
                last_use[y] = k
    live, max_live = set(), 0
    for k, (kind, x) in enumerate(order):
        if kind == "phase":
            lines.append(phase_line(x))
            continue
        if kind == "in":
            if inline_fields and x.name.startswith("f_"):
                continue
            nm = "i_" + x.name
            names[x] = nm
            lines.append(f"    const T {nm} = {input_fn(x)};")
            live.add(x)
        elif kind == "tmp":
            nm = f"t{ntmp}"
            ntmp += 1
            if pin:
                lines.append(f"    T {nm} = {pr.doprint(defs[x])};")
                lines.append(f"    DGR_PIN({nm});")
            else:
                lines.append(f"    const T {nm} = {pr.doprint(defs[x])};")
            names[x] = nm
            live.add(x)
        else:
            lines.append(f"    out.template put<{x}>({pr.doprint(outs_r[x])});  // {(out_names or VARS)[x]}")
        max_live = max(max_live, len(live))
        live = {y for y in live if last_use.get(y, -1) > k}
    return lines, ntmp, max_live


GAUGE_DOC = """//
// Gauge variants (rhs.cpp:219-265), chosen at run time: with rochester != 0
// the Rochester puncture gauge (CodeGen/bssn.py:415-455; xi = BSSN_XI), and
// sslt, the slow-start lapse term the caller forms (CodeGen/bssn.py:90-99:
// 0.6 sqrt(chi) (alpha - sqrt(chi)) exp(-t^2/800), or 0), subtracted from
// alpha's right-hand side.
"""


def generate(path, gauge=False):
    t0 = time.time()
    defs, outs_r, ops_after = eliminate(OUTPUTS_GAUGE if gauge else None)
    order = schedule(defs, outs_r)
    lines, ntmp, max_live = emit(order, defs, outs_r, input_expr)
    fn, extra = ("bssn_rhs_point_gauge", "const T *xi, const int rochester, const T sslt, ") if gauge \
        else ("bssn_rhs_point", "")
    hdr = HEADER_TOP + (GAUGE_DOC if gauge else "") + f"""// Operation count: {ops_after} ({ntmp} temporaries); at most {max_live} values
// live at once in the emitted order.
//
// In: accessor with fld<V>() (24 fields, VAR order of grDef.h), d1<V, D>()
//     (first derivatives), d2<M, C>() (second derivatives of alpha, chi,
//     beta0..2, gt0..5; components 00 01 02 11 12 22).
// Out: receives put<V>(rhs) once per variable.
#pragma once

#ifndef DGR_HD
#define DGR_HD
#endif
#ifndef DGR_INLINE
#define DGR_INLINE inline
#endif
#ifndef DGR_PIN
#define DGR_PIN(x) ((void)0)
#endif

namespace dgr {{

template <typename T, typename In, typename Out>
DGR_HD DGR_INLINE void {fn}(const In &in, const T eta, const T *lam, const T *lamf, {extra}Out &out) {{
{chr(10).join(lines)}
}}

}}  // namespace dgr
"""
    with open(path, "w") as fh:
        fh.write(hdr)
    print(f"generated {path}: {ntmp} temporaries, ops {ops_after}, max live {max_live}, "
          f"{time.time() - t0:.1f}s")


# Order in which the variables stream through the fused kernel's LDS.
# Found by local search over orders: the smallest live set of the schedule
# (126 values), and among those few consecutive phases that both need mixed
# derivatives (3 pairs), so that LDS-heavy and arithmetic-heavy phases
# alternate (measured +0.7% over an order with 6 such pairs); alpha and chi
# sit at even positions (the mixed-precision kernels with two LDS region
# buffers stage those two in FP64, and only the even-phase buffer is
# FP64-sized).
PHASE_ORDER = ["chi", "Gt1", "gt2", "gt3", "gt1", "Gt0", "gt4", "gt5", "Gt2", "gt0", "At2", "K",
               "beta2", "At3", "beta1", "B1", "alpha", "At0", "beta0", "At1", "At4", "At5", "B2", "B0"]


def _var_of(sym):
    n = sym.name
    if n.startswith("d1_") or n.startswith("d2_"):
        return n[3:].rsplit("_", 1)[0]
    return None


def schedule_phased(defs, outs, order_vars):
    """Statement order for a kernel that streams one variable at a time: the
    derivatives of variable v exist only during phase P(v), so a statement can
    run no earlier than the phase of its latest derivative input, and every
    derivative input must be placed in its own phase. Outputs run after the last
    phase (when every Kreiss-Oliger term is ready). Bottom-up over phases: place
    the statements that must sit in the current phase, plus earlier-capable ones
    that do not grow the live set, choosing each time the one adding the fewest
    live values. Fields are re-read at use and not counted as live."""
    P = {v: i for i, v in enumerate(order_vars)}
    NP = len(order_vars)
    nodes = {}

    def operands(e):
        return [("tmp", x) if x in defs else ("in", x)
                for x in e.free_symbols if x in defs or x in INPUTS]

    for s_, e in defs.items():
        nodes[("tmp", s_)] = operands(e)
    for v, e in enumerate(outs):
        nodes[("out", v)] = operands(e)
    used = {x for o in nodes.values() for x in o}
    for x in INPUTS:
        if ("in", x) in used:
            nodes[("in", x)] = []
    memo = {}

    def ep(n):
        if n not in memo:
            if n[0] == "in":
                v = _var_of(n[1])
                memo[n] = P[v] if v else 0
            else:
                r = max([ep(x) for x in nodes[n]] + [0])
                memo[n] = NP if n[0] == "out" else r
        return memo[n]

    for n in nodes:
        ep(n)
    users = {n: 0 for n in nodes}
    for n, o in nodes.items():
        for x in set(o):
            users[x] += 1

    def counted(x):
        return not (x[0] == "in" and x[1].name.startswith("f_"))

    live, rev = set(), []
    ready = {n for n in nodes if users[n] == 0}
    for p in range(NP, -1, -1):
        while True:
            best, bs = None, None
            for n in ready:
                if ep(n) > p:
                    continue
                if n[0] == "in" and _var_of(n[1]) and ep(n) != p:
                    continue  # derivatives exist only during their own phase
                new = sum(1 for x in set(nodes[n]) if x not in live and counted(x))
                gain = new - (1 if n in live and counted(n) else 0)
                if ep(n) < p and gain > 0:
                    continue
                sc = (gain, 0 if ep(n) == p else 1, -len(nodes[n]),
                      n[0], n[1] if isinstance(n[1], int) else n[1].name)
                if bs is None or sc < bs:
                    best, bs = n, sc
            if best is None:
                break
            ready.discard(best)
            rev.append(best)
            live.discard(best)
            for x in set(nodes[best]):
                live.add(x)
                users[x] -= 1
                if users[x] == 0:
                    ready.add(x)
        rev.append(("phase", p))
        assert not [n for n in ready if ep(n) == p and not (n[0] == "in" and not _var_of(n[1]))]
    assert not ready
    prog = rev[::-1]
    return prog


def generate_phased(path, gauge=False):
    t0 = time.time()
    defs, outs_r, ops_after = eliminate(OUTPUTS_GAUGE if gauge else None)
    prog = schedule_phased(defs, outs_r, PHASE_ORDER)
    vidx = [VIDX[v] for v in PHASE_ORDER]

    def phase_line(p):
        if p == len(PHASE_ORDER):
            return "    ctx.finish();"
        nxt = vidx[p + 1] if p + 1 < len(vidx) else -1
        return f"    ctx.template phase<{vidx[p]}, {nxt}, {p}>();  // {PHASE_ORDER[p]}"

    def phased_input(x):
        n = x.name
        if n.startswith("f_"):
            return f"ctx.template fld<{VIDX[n[2:]]}>()"
        return input_expr(x).replace("in.template", "ctx.template")

    lines, ntmp, _ = emit(prog, defs, outs_r, phased_input, phase_line, inline_fields=True, pin=True)
    # live set without fields (they are re-read when used)
    last_use, live, max_live = {}, set(), 0
    for k, (kind, x) in enumerate(prog):
        e = defs[x] if kind == "tmp" else outs_r[x] if kind == "out" else None
        if e is not None:
            for y in e.free_symbols:
                last_use[y] = k
    for k, (kind, x) in enumerate(prog):
        if kind in ("in", "tmp") and not x.name.startswith("f_"):
            live.add(x)
        max_live = max(max_live, len(live))
        live = {y for y in live if last_use.get(y, -1) > k}
    fn, extra = ("bssn_rhs_phased_gauge", "const T *xi, const int rochester, const T &sslt, ") if gauge \
        else ("bssn_rhs_phased", "")
    consts = "" if gauge else f"""
// first variable of the phase order (the kernel fetches it before calling)
constexpr int kBssnFirstPhaseVar = {vidx[0]};
// the whole phase order (variable of phase K)
constexpr int kBssnPhaseVar[{len(vidx)}] = {{{', '.join(str(v) for v in vidx)}}};
"""
    hdr = HEADER_TOP + (GAUGE_DOC + "// sslt is read by reference, like eta.\n" if gauge else "") + f"""//
// Phased form for a kernel that streams the variables through LDS one at a
// time, in the order {', '.join(PHASE_ORDER)}.
// ctx.phase<V, NEXT, K>() (K: position of V in the order) makes variable V's
// derivatives at this thread's point available (and may start fetching NEXT);
// ctx.d1<V, D>() / ctx.d2<M, C>() may only be called in V's phase;
// ctx.fld<V>() may be called anywhere; eta is read by reference where B's
// right-hand sides use it, after the last phase, so a phase may still set it.
// Every thread of a workgroup must execute the whole function (the phases
// synchronize the workgroup). DGR_PIN(t) marks
// where a temporary must be computed: a device build defines it to keep the
// compiler from sinking a statement into a later phase (which would keep the
// statement's inputs alive across phases).
// Operation count: {ops_after} ({ntmp} temporaries); at most {max_live} values
// (fields excluded) live at once in the emitted order.
#pragma once

#ifndef DGR_HD
#define DGR_HD
#endif
#ifndef DGR_INLINE
#define DGR_INLINE inline
#endif
#ifndef DGR_PIN
#define DGR_PIN(x) ((void)0)
#endif

namespace dgr {{
{consts}
template <typename T, typename Ctx, typename Out>
DGR_HD DGR_INLINE void {fn}(Ctx &ctx, const T &eta, const T *lam, const T *lamf, {extra}Out &out) {{
{chr(10).join(lines)}
}}

}}  // namespace dgr
"""
    with open(path, "w") as fh:
        fh.write(hdr)
    print(f"generated {path}: {ntmp} temporaries, ops {ops_after}, max live {max_live} (phased), "
          f"{time.time() - t0:.1f}s")


def generate_ham(path):
    t0 = time.time()
    defs, outs_r, ops_after = eliminate([HAM] + MOM + psi4_exprs())
    order = schedule(defs, outs_r)
    lines, ntmp, max_live = emit(order, defs, outs_r, input_expr,
                                 out_names=["ham", "mom0", "mom1", "mom2", "psi4_real", "psi4_img"])
    hdr = """// Generated by dendro-gr_amd/codegen/gen_bssn_rhs.py --ham -- do not edit.
//
// Pointwise constraints of the BSSN variables (CodeGen/constraints.py:250-282,
// which the reference expands into BSSN_GR/src/physconeqs.cpp): the
// Hamiltonian constraint chi igt^jk R_jk - At_ij At^ij + 2/3 K^2 and the
// momentum constraints igt^jk (d_k At_ij - C2^m_ki At_jm) - Gt^j At_ij
// - 3/2 igt^jk At_ki d_j chi / chi - 2/3 d_i K and the Weyl scalar psi4
// (CodeGen/constraints.py:42-244) in the tetrad built from the point's
// coordinates, from the same named intermediates (inverse metric, Christoffel
// symbols, Ricci tensor, raised At) as the right-hand side.
""" + f"""// Operation count: {ops_after} ({ntmp} temporaries); at most {max_live} values live.
//
// In: the accessor of bssn_rhs_point (fields, first derivatives of chi, K,
//     Gt, gt, At; second derivatives of chi and gt are the ones read) plus
//     crd<0..2>(), the point's coordinates (read by psi4 only).
// Out: receives put<0>(ham), put<1..3>(mom0..2), put<4>(psi4 real),
//     put<5>(psi4 imaginary). psi4 is singular on the z axis; the caller
//     replaces it there (physcon.cpp:82-85).
#pragma once

#ifndef DGR_HD
#define DGR_HD
#endif
#ifndef DGR_INLINE
#define DGR_INLINE inline
#endif

namespace dgr {{

template <typename T, typename In, typename Out>
DGR_HD DGR_INLINE void bssn_ham_point(const In &in, Out &out) {{
{chr(10).join(lines)}
}}

}}  // namespace dgr
"""
    with open(path, "w") as fh:
        fh.write(hdr)
    print(f"generated {path}: {ntmp} temporaries, ops {ops_after}, max live {max_live}, "
          f"{time.time() - t0:.1f}s")


# Phase order of the fused constraint kernel (the variables whose derivatives
# the constraints read): chosen by local search for the smallest live set of
# the schedule (see search_cons_order).
CONS_PHASE_ORDER = ["gt3", "gt1", "gt5", "gt2", "gt4", "Gt2", "Gt0", "Gt1", "gt0", "chi", "At5", "K",
                    "At4", "At0", "At2", "At3", "At1"]


def _max_live(prog, defs, outs_r):
    last_use, live, max_live = {}, set(), 0
    for k, (kind, x) in enumerate(prog):
        e = defs[x] if kind == "tmp" else outs_r[x] if kind == "out" else None
        if e is not None:
            for y in e.free_symbols:
                last_use[y] = k
    for k, (kind, x) in enumerate(prog):
        if kind in ("in", "tmp") and not x.name.startswith("f_"):
            live.add(x)
        max_live = max(max_live, len(live))
        live = {y for y in live if last_use.get(y, -1) > k}
    return max_live


def search_cons_order(defs, outs_r, start, rounds=60, seed=1):
    """Local search over phase orders (swaps of two positions), keeping a swap
    when the live set of the phased schedule does not grow."""
    import random
    rnd = random.Random(seed)
    best = list(start)
    best_live = _max_live(schedule_phased(defs, outs_r, best), defs, outs_r)
    for _ in range(rounds):
        cand = list(best)
        i, j = rnd.sample(range(len(cand)), 2)
        cand[i], cand[j] = cand[j], cand[i]
        lv = _max_live(schedule_phased(defs, outs_r, cand), defs, outs_r)
        if lv <= best_live:
            best, best_live = cand, lv
    return best, best_live


def generate_cons_phased(path, search=False):
    t0 = time.time()
    defs, outs_r, ops_after = eliminate([HAM] + MOM + psi4_exprs())
    used = {_var_of(x) for e in list(defs.values()) + outs_r for x in e.free_symbols if _var_of(x)}
    order = [v for v in CONS_PHASE_ORDER if v in used]
    assert set(order) == used, (sorted(used), order)
    if search:
        order, lv = search_cons_order(defs, outs_r, order)
        print(f"phase order {order}: max live {lv}")
    prog = schedule_phased(defs, outs_r, order)
    vidx = [VIDX[v] for v in order]

    def phase_line(p):
        if p == len(order):
            return "    ctx.finish();"
        nxt = vidx[p + 1] if p + 1 < len(vidx) else -1
        return f"    ctx.template phase<{vidx[p]}, {nxt}, {p}>();  // {order[p]}"

    def phased_input(x):
        n = x.name
        if n.startswith("f_"):
            return f"ctx.template fld<{VIDX[n[2:]]}>()"
        return input_expr(x).replace("in.template", "ctx.template")

    lines, ntmp, _ = emit(prog, defs, outs_r, phased_input, phase_line, inline_fields=True, pin=True,
                          out_names=["ham", "mom0", "mom1", "mom2", "psi4_real", "psi4_img"])
    max_live = _max_live(prog, defs, outs_r)
    hdr = """// Generated by dendro-gr_amd/codegen/gen_bssn_rhs.py --cons-phased -- do not edit.
//
// The constraints of bssn_ham_generated.h (CodeGen/constraints.py:42-282:
// Hamiltonian, momentum, psi4) in the phased form of bssn_rhs_phased.h: the
// variables whose derivatives they read stream through the kernel one at a
// time, in the order """ + ", ".join(order) + f""".
// ctx.phase<V, NEXT, K>() makes variable V's derivatives available at this
// thread's point; ctx.d1 / ctx.d2 only in V's phase; ctx.fld<V>() and
// ctx.crd<D>() (the point's coordinates, read by psi4) anywhere. Every thread
// of a workgroup must execute the whole function. Outputs: put<0>(ham),
// put<1..3>(mom0..2), put<4..5>(psi4 real, imaginary).
// Operation count: {ops_after} ({ntmp} temporaries); at most {max_live} values
// (fields excluded) live at once in the emitted order.
#pragma once

#ifndef DGR_HD
#define DGR_HD
#endif
#ifndef DGR_INLINE
#define DGR_INLINE inline
#endif
#ifndef DGR_PIN
#define DGR_PIN(x) ((void)0)
#endif

namespace dgr {{

// phases of the constraint kernel and the variable of each
constexpr int kConsPhases = {len(vidx)};
constexpr int kConsPhaseVar[{len(vidx)}] = {{{', '.join(str(v) for v in vidx)}}};

template <typename T, typename Ctx, typename Out>
DGR_HD DGR_INLINE void bssn_cons_phased(Ctx &ctx, Out &out) {{
{chr(10).join(lines)}
}}

}}  // namespace dgr
"""
    with open(path, "w") as fh:
        fh.write(hdr)
    print(f"generated {path}: {ntmp} temporaries, ops {ops_after}, max live {max_live} (phased), "
          f"{time.time() - t0:.1f}s")


CSE_MODE = "local"

if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    for a in sys.argv[1:]:
        if a.startswith("--cse="):
            CSE_MODE = a.split("=", 1)[1]
    gauge = "--gauge" in sys.argv
    if "--phased" in sys.argv:
        generate_phased(args[0] if args else "bssn_rhs_phased.h", gauge)
    elif "--cons-phased" in sys.argv:
        generate_cons_phased(args[0] if args else "bssn_cons_phased.h", search="--search" in sys.argv)
    elif "--ham" in sys.argv:
        generate_ham(args[0] if args else "bssn_ham_generated.h")
    else:
        generate(args[0] if args else "bssn_rhs_generated.h", gauge)
