"""Generator of the gfx950 one-electron kernels (S, T, V).

One kernel pair per angular-momentum pair (la, lb), la, lb in {0, 1, 2}
(L shells enter as their s and p parts, weighted by the p/s coefficient
ratio, spread_basis.f90:149-151).  One thread per work item (shell a,
shell b, part a, part b).  Obara--Saika for Cartesian Gaussians:

* overlap, per axis  s(i+1,j) = PA s(i,j) + (i s(i-1,j) + j s(i,j-1)) / 2p,
                     s(0,j+1) = PB s(0,j) + j s(0,j-1) / 2p;
* kinetic, per axis  t(i,j) = b(2j+1) s(i,j) - 2b^2 s(i,j+2) - j(j-1)/2 s(i,j-2),
  T = tx sy sz + sx ty sz + sx sy tz;
* nuclear attraction on the first centre with the auxiliary index m,
  [e+1_i]^(m) = PA_i [e]^(m) - PC_i [e]^(m+1) + e_i/2p ([e-1_i]^(m) - [e-1_i]^(m+1)),
  [0]^(m) = -Z 2pi/p exp(-ab/p AB^2) F_m(p PC^2), summed over primitives and
  nuclei, then moved to the second centre by (a|b+1_i) = (a+1_i|b) + AB_i (a|b)
  (the operator does not depend on A or B);

then d to the five real spherical functions (index_functions.f90:147-168).
The reference computes the same matrices in cmpt1e.f90:1-258 with
one_electron_terms.f90:7-1058.

    python -m hfcode_amd.gen.gen_onee OUTDIR   (the build does this)
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from gen_eri import AX, SPH, add1, cart, sub1, tname  # noqa: E402


def ncart(l):
    return (l + 1) * (l + 2) // 2


def sph_terms(l, k):
    """(Cartesian triple, factor) list of AO component k of a part of angular momentum l."""
    if l == 2:
        return SPH[k]
    return [(cart(l)[k], "1")]


def ncomp(l):
    return 5 if l == 2 else ncart(l)


def emit_common(W, la, lb):
    W("    const int4 w = a.items[it];")
    W("    const int sa = w.x, sb = w.y, pa = w.z, pb = w.w;")
    W("    const double Ax = a.sh_xyz[3 * sa], Ay = a.sh_xyz[3 * sa + 1], Az = a.sh_xyz[3 * sa + 2];")
    W("    const double Bx = a.sh_xyz[3 * sb], By = a.sh_xyz[3 * sb + 1], Bz = a.sh_xyz[3 * sb + 2];")
    W("    const double ABx = Ax - Bx, ABy = Ay - By, ABz = Az - Bz;")
    W("    const double AB2 = ABx * ABx + ABy * ABy + ABz * ABz;")
    W("    const int npa = a.sh_np[sa], npb = a.sh_np[sb];")
    W("    const bool wla = a.sh_type[sa] == 12 && pa == 1, wlb = a.sh_type[sb] == 12 && pb == 1;")


def emit_prim_head(W):
    W("        const double za = a.sh_z[sa * a.maxprim + ka], zb = a.sh_z[sb * a.maxprim + kb];")
    W("        const double p = za + zb, oo2p = 0.5 / p;")
    W("        double wgt = a.sh_c[sa * a.maxprim + ka] * a.sh_c[sb * a.maxprim + kb];")
    W("        if (wla) wgt *= a.sh_cl[sa * a.maxprim + ka];")
    W("        if (wlb) wgt *= a.sh_cl[sb * a.maxprim + kb];")
    W("        const double Kx = exp(-za * zb / p * AB2);")
    W("        const double Px = (za * Ax + zb * Bx) / p, Py = (za * Ay + zb * By) / p, Pz = (za * Az + zb * Bz) / p;")
    W("        const double PAx = Px - Ax, PAy = Py - Ay, PAz = Pz - Az;")
    W("        const double PBx = Px - Bx, PBy = Py - By, PBz = Pz - Bz;")


def emit_store(W, la, lb, mat, val):
    """Spherical transform of the Cartesian block val(x, y) and the two stores."""
    for ka in range(ncomp(la)):
        for kb in range(ncomp(lb)):
            terms = []
            for (ea, fa) in sph_terms(la, ka):
                for (eb, fb) in sph_terms(lb, kb):
                    terms.append((val(ea, eb), fa, fb))
            ex = " + ".join(t if (fa, fb) == ("1", "1") else "(%s * %s) * %s" % (fa, fb, t) for (t, fa, fb) in terms)
            W("    { const double v = %s;" % ex)
            W("      const int mu = a.sh_ao[sa] + (pa == 1 ? 1 : 0) + %d, nu = a.sh_ao[sb] + (pb == 1 ? 1 : 0) + %d;" % (ka, kb))
            W("      a.%s[(size_t)mu * a.nb + nu] = v; a.%s[(size_t)nu * a.nb + mu] = v; }" % (mat, mat))


def kernel_st(W, la, lb):
    W("__global__ void __launch_bounds__(HFX_NT) k_onee_st_%d%d(const OneeArgs a) {" % (la, lb))
    W("  const int stride = gridDim.x * blockDim.x;")
    W("  for (int it = blockIdx.x * blockDim.x + threadIdx.x; it < a.nitems; it += stride) {")
    emit_common(W, la, lb)
    pairs = [(ea, eb) for ea in cart(la) for eb in cart(lb)]
    for ea, eb in pairs:
        W("    double s_%s_%s = 0, t_%s_%s = 0;" % (tname(ea), tname(eb), tname(ea), tname(eb)))
    W("    for (int ka = 0; ka < npa; ++ka) {")
    W("      for (int kb = 0; kb < npb; ++kb) {")
    emit_prim_head(W)
    W("        const double s00 = wgt * Kx * pow(3.141592653589793 / p, 1.5);")
    imax, jmax = la, lb + 2
    for k in range(3):
        X = AX[k]
        done = set()
        W("        const double s%s_0_0 = 1.0;" % X)
        for i in range(imax + 1):
            for j in range(jmax + 1):
                if i == 0 and j == 0:
                    continue
                if i > 0:
                    ex = "PA%s * s%s_%d_%d" % (X, X, i - 1, j)
                    if i > 1:
                        ex += " + %d * oo2p * s%s_%d_%d" % (i - 1, X, i - 2, j)
                    if j > 0:
                        ex += " + %d * oo2p * s%s_%d_%d" % (j, X, i - 1, j - 1)
                else:
                    ex = "PB%s * s%s_%d_%d" % (X, X, i, j - 1)
                    if j > 1:
                        ex += " + %d * oo2p * s%s_%d_%d" % (j - 1, X, i, j - 2)
                W("        const double s%s_%d_%d = %s;" % (X, i, j, ex))
    for ea, eb in pairs:
        sx = ["s%s_%d_%d" % (AX[k], ea[k], eb[k]) for k in range(3)]
        tx = []
        for k in range(3):
            i, j = ea[k], eb[k]
            t = "zb * %d * s%s_%d_%d - 2.0 * zb * zb * s%s_%d_%d" % (2 * j + 1, AX[k], i, j, AX[k], i, j + 2)
            if j >= 2:
                t += " - %s * s%s_%d_%d" % (repr(0.5 * j * (j - 1)), AX[k], i, j - 2)
            tx.append("(%s)" % t)
        W("        s_%s_%s += s00 * %s * %s * %s;" % (tname(ea), tname(eb), sx[0], sx[1], sx[2]))
        W("        t_%s_%s += s00 * (%s * %s * %s + %s * %s * %s + %s * %s * %s);" % (
            tname(ea), tname(eb), tx[0], sx[1], sx[2], sx[0], tx[1], sx[2], sx[0], sx[1], tx[2]))
    W("      }")
    W("    }")
    emit_store(W, la, lb, "S", lambda ea, eb: "s_%s_%s" % (tname(ea), tname(eb)))
    emit_store(W, la, lb, "T", lambda ea, eb: "t_%s_%s" % (tname(ea), tname(eb)))
    W("  }")
    W("}")


def kernel_v(W, la, lb):
    M = la + lb
    W("__global__ void __launch_bounds__(HFX_NT) k_onee_v_%d%d(const OneeArgs a) {" % (la, lb))
    W("  extern __shared__ char hfx_smem[];")
    W("  double* btab = reinterpret_cast<double*>(hfx_smem);")
    W("  constexpr int NB_COL = %d + BoysTerms<double>::K;" % M)
    W("  for (int i = threadIdx.x; i < HFX_BOYS_ROWS * NB_COL; i += blockDim.x) {")
    W("    const int r = i / NB_COL, col = i - r * NB_COL;")
    W("    btab[i] = a.boys64[r * HFX_BOYS_COLS + col];")
    W("  }")
    W("  __syncthreads();")
    W("  const int stride = gridDim.x * blockDim.x;")
    W("  for (int it = blockIdx.x * blockDim.x + threadIdx.x; it < a.nitems; it += stride) {")
    emit_common(W, la, lb)
    targets = [e for l in range(la, la + lb + 1) for e in cart(l)]
    for e in targets:
        W("    double ve_%s = 0;" % tname(e))
    W("    for (int ka = 0; ka < npa; ++ka) {")
    W("      for (int kb = 0; kb < npb; ++kb) {")
    emit_prim_head(W)
    W("        const double pref = wgt * Kx * (6.283185307179586 / p);")
    W("        for (int n = 0; n < a.natom; ++n) {")
    W("          const double PCx = Px - a.coords[3 * n], PCy = Py - a.coords[3 * n + 1], PCz = Pz - a.coords[3 * n + 2];")
    W("          double F[%d];" % (M + 1))
    W("          boys<double, %d>(btab, NB_COL, p * (PCx * PCx + PCy * PCy + PCz * PCz), F);" % M)
    W("          const double q = -a.charges[n] * pref;")
    memo = {}

    def get(e, m):
        if min(e) < 0:
            return None
        key = (e, m)
        if key in memo:
            return memo[key]
        if sum(e) == 0:
            nm = "v000_%d" % m
            W("          const double %s = q * F[%d];" % (nm, m))
        else:
            i = next(k for k in range(3) if e[k] > 0)
            e1 = sub1(e, i)
            t1, t2 = get(e1, m), get(e1, m + 1)
            ex = "PA%s * %s - PC%s * %s" % (AX[i], t1, AX[i], t2)
            if e1[i] > 0:
                e2 = sub1(e1, i)
                t3, t4 = get(e2, m), get(e2, m + 1)
                ex += " + %d * oo2p * (%s - %s)" % (e1[i], t3, t4)
            nm = "v%s_%d" % (tname(e), m)
            W("          const double %s = %s;" % (nm, ex))
        memo[key] = nm
        return nm

    for e in targets:
        W("          ve_%s += %s;" % (tname(e), get(e, 0)))
    W("        }")
    W("      }")
    W("    }")
    hm = {}

    def hb(ea, eb):
        key = (ea, eb)
        if key in hm:
            return hm[key]
        if sum(eb) == 0:
            hm[key] = "ve_%s" % tname(ea)
            return hm[key]
        i = next(k for k in range(3) if eb[k] > 0)
        x1, x2 = hb(add1(ea, i), sub1(eb, i)), hb(ea, sub1(eb, i))
        nm = "h_%s_%s" % (tname(ea), tname(eb))
        W("    const double %s = %s + AB%s * %s;" % (nm, x1, AX[i], x2))
        hm[key] = nm
        return nm

    emit_store(W, la, lb, "V", hb)
    W("  }")
    W("}")


def source():
    L = []
    W = L.append
    W("// One-electron kernels, generated by hfcode_amd/gen/gen_onee.py -- do not edit")
    W("#include \"hfx_device.h\"")
    W("#include \"hfx_kernels.h\"")
    W("namespace hfx {")
    for la in range(3):
        for lb in range(3):
            kernel_st(W, la, lb)
            kernel_v(W, la, lb)
    W("const OneeKernel kOneeST[3][3] = {")
    W("  " + ", ".join("{%s}" % ", ".join("k_onee_st_%d%d" % (la, lb) for lb in range(3)) for la in range(3)))
    W("};")
    W("const OneeKernel kOneeV[3][3] = {")
    W("  " + ", ".join("{%s}" % ", ".join("k_onee_v_%d%d" % (la, lb) for lb in range(3)) for la in range(3)))
    W("};")
    W("}  // namespace hfx")
    return "\n".join(L) + "\n"


def main(outdir):
    os.makedirs(outdir, exist_ok=True)
    path = os.path.join(outdir, "onee.hip")
    src = source()
    if not os.path.exists(path) or open(path).read() != src:
        with open(path, "w") as fh:
            fh.write(src)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(HERE), "csrc", "gen"))
