"""Generator of the per-class gfx950 ERI + Fock-digestion kernels.

One HIP translation unit per shell-quartet class (ab|cd), a, b, c, d in
{s, p, d, l}.  The 46 classes and their order are those of the class loops of
fock_build_non_diag.f90 (:52 (ss|ss) ... :3289 (ll|ll)); the integrals are
the ones the reference's i2e_* routines return (s_sp_l_terms.f90,
s_p_d_terms.f90, s_d_l_terms.f90).  The algorithm is written from the
Head-Gordon--Pople / Obara--Saika equations, not from those routines:

1. primitive loop over bra x ket primitive pairs:
   [00|00]^(m) = K_ab K_cd (p+q)^-1/2 F_m(T), Boys F_m from an LDS table;
   vertical recursion on A, then on C (HGP eqs. 6 and its ket twin) up to
   [e0|f0]^(0), e in [la, la+lb], f in [lc, lc+ld];
2. contraction, with the L-shell p/s coefficient ratio (djkl,
   spread_basis.f90:149-151) applied to the p parts of L shells;
3. horizontal recursion (a,b+1_i| = (a+1_i,b| + AB_i (a,b| on bra then ket;
4. Cartesian -> spherical d (z2, xz, yz, xy, x2-y2; index_functions.f90:147-168);
5. streamed digestion: every element updates six partial J/K contractions in
   FP64; the quartet is dropped if max|e| < thr (fock_build_non_diag.f90:69)
   and otherwise added to the accumulator X with atomics (F += X + X^T).

Within a shell pair the shell of higher angular momentum is the build centre
(no HRR for (ds|, (ls|, ...); the digestion works on AO indices, so the
orientation does not change any result.

Run as ``python -m hfcode_amd.gen.gen_eri OUTDIR`` (the build does this).
It also writes ``flops_per_class.json``: operation counts of the emitted
code (add = mul = 1, fma = 2), per primitive quartet, per bra primitive and
per contracted quartet, which price the algorithmic FLOPs of a build.
"""
import json
import os
import sys

CLASSES = ("ssss sssp sssd sssl sspp sspd sspl ssdd ssdl ssll spsp spsd spsl sppp sppd "
           "sppl spdd spll sdsd sdsl sdpp sdpd sddd sddl sdll slsl slpp slpl sldd sldl "
           "slll pppp pppd pppl ppdd ppll pdpd pddd plpl plll dddd dddl ddll dldl dlll llll").split()
PAIR_TYPES = ["ss", "sp", "sd", "sl", "pp", "pd", "pl", "dd", "dl", "ll"]

# angular parts of a shell: (l, weighted-by-djkl)
PARTS = {"s": [(0, False)], "p": [(1, False)], "d": [(2, False)], "l": [(0, False), (1, True)]}
NCOMP = {"s": 1, "p": 3, "d": 5, "l": 4}
ORIENT_RANK = {"s": 0, "l": 1, "p": 2, "d": 3}
AX = "xyz"
SQ3 = "1.7320508075688772"
HSQ3 = "0.8660254037844386"
BOYS_TERMS = {"double": 8, "float": 5}
BOYS_ROWS = 449          # grid 0, 1/16, ..., 28


def cart(l):
    """Cartesian exponent triples of angular momentum l, in the reference's
    order (index_functions.f90:57-68): x^l first, z^l last."""
    out = []
    for i in range(l + 1):
        lx = l - i
        for n in range(i + 1):
            out.append((lx, i - n, n))
    return out


def tname(t):
    return "%d%d%d" % t


def sub1(t, i):
    return tuple(v - (1 if k == i else 0) for k, v in enumerate(t))


def add1(t, i):
    return tuple(v + (1 if k == i else 0) for k, v in enumerate(t))


def comps(sh):
    """Per AO component of shell type `sh`: (part index, kind, data).
    kind 'c' -> Cartesian triple, kind 'sph' -> spherical d index 0..4."""
    if sh == "s":
        return [(0, "c", (0, 0, 0))]
    if sh == "p":
        return [(0, "c", t) for t in cart(1)]
    if sh == "d":
        return [(0, "sph", k) for k in range(5)]
    return [(0, "c", (0, 0, 0))] + [(1, "c", t) for t in cart(1)]


# spherical d from Cartesian d (order xx, xy, xz, yy, yz, zz), matharmond
# (index_functions.f90:151-166)
SPH = [
    [((0, 0, 2), "1"), ((2, 0, 0), "-0.5"), ((0, 2, 0), "-0.5")],
    [((1, 0, 1), SQ3)],
    [((0, 1, 1), SQ3)],
    [((1, 1, 0), SQ3)],
    [((2, 0, 0), HSQ3), ((0, 2, 0), "-" + HSQ3)],
]


class Ops:
    def __init__(self):
        self.n = 0


class Gen:
    def __init__(self, cls):
        self.cls = cls
        ta, tb, tc, td = cls
        self.bra_type = ta + tb
        self.ket_type = tc + td
        self.swap_bra = ORIENT_RANK[tb] > ORIENT_RANK[ta]
        self.swap_ket = ORIENT_RANK[td] > ORIENT_RANK[tc]
        self.A, self.B = (tb, ta) if self.swap_bra else (ta, tb)
        self.C, self.D = (td, tc) if self.swap_ket else (tc, td)
        self.bra_parts = [(pa, pb) for pa in range(len(PARTS[self.A])) for pb in range(len(PARTS[self.B]))]
        self.ket_parts = [(pc, pd) for pc in range(len(PARTS[self.C])) for pd in range(len(PARTS[self.D]))]
        self.has_L_bra = "l" in (self.A, self.B)
        self.has_L_ket = "l" in (self.C, self.D)
        la_max = max(l for l, _ in PARTS[self.A]) + max(l for l, _ in PARTS[self.B])
        lc_max = max(l for l, _ in PARTS[self.C]) + max(l for l, _ in PARTS[self.D])
        self.M = la_max + lc_max
        # (sl|sl) is the one class routine that does not skip negligible
        # primitive pairs (i2e_slsl, s_d_l_terms.f90:1629-1693).
        self.all_prims = (cls == "slsl")

    # ---- ranges -----------------------------------------------------------
    def erange(self, bp):
        la = PARTS[self.A][bp[0]][0]
        lb = PARTS[self.B][bp[1]][0]
        return la, lb, list(range(la, la + lb + 1))

    def frange(self, kp):
        lc = PARTS[self.C][kp[0]][0]
        ld = PARTS[self.D][kp[1]][0]
        return lc, ld, list(range(lc, lc + ld + 1))

    # ---- primitive-level VRR ---------------------------------------------
    def gen_vrr(self, lines, ops, targets):
        memo = {}

        def get(e, f, m):
            if min(e) < 0 or min(f) < 0:
                return None
            key = (e, f, m)
            if key in memo:
                return memo[key]
            if sum(e) == 0 and sum(f) == 0:
                memo[key] = "s%d" % m
                return memo[key]
            name = "v%s_%s_%d" % (tname(e), tname(f), m)
            if sum(f) == 0:
                i = next(k for k in range(3) if e[k] > 0)
                e1 = sub1(e, i)
                t1 = get(e1, f, m); t2 = get(e1, f, m + 1)
                expr = "PA%s * %s + WP%s * %s" % (AX[i], t1, AX[i], t2)
                ops.n += 3
                k = e[i] - 1
                if k > 0:
                    e2 = sub1(e1, i)
                    t3 = get(e2, f, m); t4 = get(e2, f, m + 1)
                    expr += " + (R)%d * oo2p * (%s - rho_p * %s)" % (k, t3, t4)
                    ops.n += 4
            else:
                i = next(k for k in range(3) if f[k] > 0)
                f1 = sub1(f, i)
                t1 = get(e, f1, m); t2 = get(e, f1, m + 1)
                expr = "QC%s * %s + WQ%s * %s" % (AX[i], t1, AX[i], t2)
                ops.n += 3
                k = f[i] - 1
                if k > 0:
                    f2 = sub1(f1, i)
                    t3 = get(e, f2, m); t4 = get(e, f2, m + 1)
                    expr += " + (R)%d * oo2q * (%s - rho_q * %s)" % (k, t3, t4)
                    ops.n += 4
                if e[i] > 0:
                    t5 = get(sub1(e, i), f1, m + 1)
                    expr += " + (R)%d * oo2pq * %s" % (e[i], t5)
                    ops.n += 2
            lines.append("const R %s = %s;" % (name, expr))
            memo[key] = name
            return name

        out = {}
        for (e, f) in targets:
            out[(e, f)] = get(e, f, 0)
        return out

    # ---- code emission ----------------------------------------------------
    def weights(self):
        def wket(kp):
            t = []
            if PARTS[self.C][kp[0]][1]:
                t.append("wC")
            if PARTS[self.D][kp[1]][1]:
                t.append("wD")
            return "*".join(t)

        def wbra(bp):
            t = []
            if PARTS[self.A][bp[0]][1]:
                t.append("wA")
            if PARTS[self.B][bp[1]][1]:
                t.append("wB")
            return "*".join(t)
        return wbra, wket

    def emit_quartet_setup(self, W):
        oa = "b" if self.swap_bra else "a"
        ob = "a" if self.swap_bra else "b"
        oc = "b" if self.swap_ket else "a"
        od = "a" if self.swap_ket else "b"
        W("    const PairRec& pb = a.pairs[bk.x];")
        W("    const PairRec& pk = a.pairs[bk.y];")
        W("    const double Ax = pb.%sx, Ay = pb.%sy, Az = pb.%sz;" % (oa, oa, oa))
        W("    const double Cx = pk.%sx, Cy = pk.%sy, Cz = pk.%sz;" % (oc, oc, oc))
        W("    const R ABx = (R)(Ax - pb.%sx), ABy = (R)(Ay - pb.%sy), ABz = (R)(Az - pb.%sz);" % (ob, ob, ob))
        W("    const R CDx = (R)(Cx - pk.%sx), CDy = (R)(Cy - pk.%sy), CDz = (R)(Cz - pk.%sz);" % (od, od, od))
        W("    (void)ABx; (void)ABy; (void)ABz; (void)CDx; (void)CDy; (void)CDz;")
        # FP32 primitive loops: P - Q = (A - C) + (P - A) - (Q - C), the centres' difference from FP64
        W("    const R ACx = (R)(Ax - Cx), ACy = (R)(Ay - Cy), ACz = (R)(Az - Cz); (void)ACx; (void)ACy; (void)ACz;")
        W("    const int shA = pb.sh%s, shB = pb.sh%s, shC = pk.sh%s, shD = pk.sh%s;" % (oa, ob, oc, od))
        W("    const int aoA = pb.ao%s, aoB = pb.ao%s, aoC = pk.ao%s, aoD = pk.ao%s;" % (oa, ob, oc, od))
        W("    (void)shA; (void)shB; (void)shC; (void)shD; (void)aoA; (void)aoB; (void)aoC; (void)aoD;")
        # the quartet's molecule: its P / X block starts at mbase (stacked nb x nb matrices
        # of a batch context; 0 with one molecule) and its threshold may be its own
        W("    const long long mq = pb.mbase; (void)mq;")

    def emit_group(self, W, combos, ops_prim, ops_bra, ops_ctr, gid, sink=None, fsel=None, hrr="ket_first",
                   ext=None):
        """Primitive loops + contraction + HRR + spherical transform for the
        part combinations `combos`; returns {(ia,ib,ic,id): element var}.
        With `sink` (block kernels) the HRR runs ket first, then the bra per
        ket component, and sink(ia, ib, ic, id, var) is called as soon as an
        element exists, so that it can be stored and dropped at once.
        Split combinations (one part combination whose accumulators would not
        fit the registers): fsel restricts the contracted targets [e0|f0] to
        the ket components f in fsel; hrr="acc" then returns
        {(e, f): accumulator} without an HRR, and hrr="cd_loop" runs the HRR
        one ket output at a time, taking [e0|f0] for f outside fsel from
        ext(e, f) (an expression, e.g. a load of what another kernel stored)."""
        wbra, wket = self.weights()
        bps = sorted({c[0] for c in combos})
        kps_by_bp = {bp: [c[1] for c in combos if c[0] == bp] for bp in bps}
        kps = sorted({c[1] for c in combos})
        e_by_bp = {bp: self.erange(bp)[2] for bp in bps}
        f_by_kp = {kp: self.frange(kp)[2] for kp in kps}
        e_all = sorted({l for bp in bps for l in e_by_bp[bp]})
        # several bra parts: contract the ket sum first (tmp), then weight by
        # the bra part; one bra part: its weight folds into the ket weight
        two_level = len(bps) > 1
        M = max(max(e_by_bp[bp]) for bp in bps) + max(max(f_by_kp[kp]) for kp in kps)
        targets = []
        seen = set()
        def fl(kp):   # the ket components f of part kp that this kernel contracts
            return [f for lf in f_by_kp[kp] for f in cart(lf) if fsel is None or f in fsel]

        for kp in kps:
            es = e_all if two_level else e_by_bp[bps[0]]
            for le in es:
                for e in cart(le):
                    for f in fl(kp):
                        if (e, f) not in seen:
                            seen.add((e, f))
                            targets.append((e, f))
        if fsel is not None:
            assert len(combos) == 1
            M = max(sum(e) + sum(f) for e, f in targets)

        def acc_name(bp, kp, e, f):
            return "acc%s_%d%d_%d%d_%s_%s" % (gid, bp[0], bp[1], kp[0], kp[1], tname(e), tname(f))

        def tmp_name(kp, e, f):
            return "tmp%s_%d%d_%s_%s" % (gid, kp[0], kp[1], tname(e), tname(f))

        for (bp, kp) in combos:
            for le in e_by_bp[bp]:
                for e in cart(le):
                    for f in fl(kp):
                        W("    R %s = 0;" % acc_name(bp, kp, e, f))
        wa_f = "wb" if self.swap_bra else "wa"
        wb_f = "wa" if self.swap_bra else "wb"
        wc_f = "wb" if self.swap_ket else "wa"
        wd_f = "wa" if self.swap_ket else "wb"
        npb = "pb.npall" if self.all_prims else "pb.np"
        npk = "pk.npall" if self.all_prims else "pk.np"
        # primitive records: FP64 with absolute centres, FP32 (prims32) with
        # the centre relative to each shell of the pair
        pa_b = "PB" if self.swap_bra else "PA"
        pa_k = "PB" if self.swap_ket else "PA"
        W("    for (int ib = 0; ib < %s; ++ib) {" % npb)
        W("      R p, pinv, Kab, PAx, PAy, PAz, wA = 1, wB = 1; (void)wA; (void)wB;")
        W("      double Pxd = 0, Pyd = 0, Pzd = 0; (void)Pxd; (void)Pyd; (void)Pzd;")
        W("      if constexpr (sizeof(R) == sizeof(double)) {")
        W("        const PrimRec& rb = a.prims[pb.p0 + ib];")
        W("        p = rb.p; pinv = rb.pinv; Kab = rb.K;")
        W("        Pxd = rb.Px; Pyd = rb.Py; Pzd = rb.Pz;")
        W("        PAx = (R)(Pxd - Ax); PAy = (R)(Pyd - Ay); PAz = (R)(Pzd - Az);")
        if self.has_L_bra:
            W("        wA = (R)rb.%s; wB = (R)rb.%s;" % (wa_f, wb_f))
        W("      } else {")
        W("        const PrimRec32& rb = a.prims32[pb.p0 + ib];")
        W("        p = rb.p; pinv = rb.pinv; Kab = rb.K;")
        W("        PAx = rb.%sx; PAy = rb.%sy; PAz = rb.%sz;" % (pa_b, pa_b, pa_b))
        if self.has_L_bra:
            W("        wA = rb.%s; wB = rb.%s;" % (wa_f, wb_f))
        W("      }")
        W("      const R oo2p = (R)0.5 * pinv; (void)oo2p;")
        ops_bra.n += 4
        if two_level:
            for kp in kps:
                for le in e_all:
                    for lf in f_by_kp[kp]:
                        for e in cart(le):
                            for f in cart(lf):
                                W("      R %s = 0;" % tmp_name(kp, e, f))
        W("      for (int ik = 0; ik < %s; ++ik) {" % npk)
        W("        R q, qinv, Kcd, QCx, QCy, QCz, PQx, PQy, PQz, wC = 1, wD = 1; (void)wC; (void)wD;")
        W("        if constexpr (sizeof(R) == sizeof(double)) {")
        W("          const PrimRec& rk = a.prims[pk.p0 + ik];")
        W("          q = rk.p; qinv = rk.pinv; Kcd = rk.K;")
        W("          QCx = (R)(rk.Px - Cx); QCy = (R)(rk.Py - Cy); QCz = (R)(rk.Pz - Cz);")
        W("          PQx = (R)(Pxd - rk.Px); PQy = (R)(Pyd - rk.Py); PQz = (R)(Pzd - rk.Pz);")
        if self.has_L_ket:
            W("          wC = (R)rk.%s; wD = (R)rk.%s;" % (wc_f, wd_f))
        W("        } else {")
        W("          const PrimRec32& rk = a.prims32[pk.p0 + ik];")
        W("          q = rk.p; qinv = rk.pinv; Kcd = rk.K;")
        W("          QCx = rk.%sx; QCy = rk.%sy; QCz = rk.%sz;" % (pa_k, pa_k, pa_k))
        W("          PQx = (ACx + PAx) - QCx; PQy = (ACy + PAy) - QCy; PQz = (ACz + PAz) - QCz;")
        if self.has_L_ket:
            W("          wC = rk.%s; wD = rk.%s;" % (wc_f, wd_f))
        W("        }")
        W("        const R ipq = hfx_rcp(p + q);")
        W("        const R rho_p = q * ipq, rho_q = p * ipq; (void)rho_p; (void)rho_q;")
        W("        const R T = p * rho_p * (PQx * PQx + PQy * PQy + PQz * PQz);")
        W("        const R oo2q = (R)0.5 * qinv, oo2pq = (R)0.5 * ipq; (void)oo2q; (void)oo2pq;")
        W("        const R WPx = -rho_p * PQx, WPy = -rho_p * PQy, WPz = -rho_p * PQz;")
        W("        const R WQx = rho_q * PQx, WQy = rho_q * PQy, WQz = rho_q * PQz;")
        W("        (void)WPx; (void)WPy; (void)WPz; (void)WQx; (void)WQy; (void)WQz;")
        W("        (void)QCx; (void)QCy; (void)QCz; (void)PAx; (void)PAy; (void)PAz;")
        ops_prim.n += 1 + 1 + 2 + 6 + 2 + 6 + 6
        W("        R F[%d];" % (M + 1))
        W("        boys<R, %d>(btab, NB_COL, T, F);" % M)
        W("        const R pref = Kab * Kcd * hfx_sqrt_rcp(p + q, ipq);")
        ops_prim.n += 3
        for m in range(M + 1):
            W("        const R s%d = pref * F[%d];" % (m, m))
        ops_prim.n += M + 1
        vlines = []
        vops = Ops()
        vmap = self.gen_vrr(vlines, vops, targets)
        ops_prim.n += vops.n
        # accumulation statements per VRR variable, emitted right after the
        # variable's definition so that a target needs no register beyond its
        # last use in the recursion
        accum = {}
        for kp in kps:
            wk = wket(kp)
            if two_level:
                for le in e_all:
                    for lf in f_by_kp[kp]:
                        for e in cart(le):
                            for f in cart(lf):
                                v = vmap[(e, f)]
                                if wk:
                                    st = "%s += %s * %s;" % (tmp_name(kp, e, f), wk, v); ops_prim.n += 2
                                else:
                                    st = "%s += %s;" % (tmp_name(kp, e, f), v); ops_prim.n += 1
                                accum.setdefault(v, []).append(st)
            else:
                bp = bps[0]
                if kp not in kps_by_bp[bp]:
                    continue
                wk = "*".join(x for x in (wbra(bp), wk) if x)
                for le in e_by_bp[bp]:
                    for e in cart(le):
                        for f in fl(kp):
                            v = vmap[(e, f)]
                            if wk:
                                st = "%s += %s * %s;" % (acc_name(bp, kp, e, f), wk, v); ops_prim.n += 2
                            else:
                                st = "%s += %s;" % (acc_name(bp, kp, e, f), v); ops_prim.n += 1
                            accum.setdefault(v, []).append(st)
        defined = {ln.split()[2] for ln in vlines}
        for v in list(accum):
            if v not in defined:           # a Boys term s<m> itself
                for st in accum.pop(v):
                    W("        " + st)
        for ln in vlines:
            W("        " + ln)
            for st in accum.pop(ln.split()[2], []):
                W("        " + st)
        assert not accum
        W("      }")
        if two_level:
            for (bp, kp) in combos:
                wb = wbra(bp)
                for le in e_by_bp[bp]:
                    for lf in f_by_kp[kp]:
                        for e in cart(le):
                            for f in cart(lf):
                                t = tmp_name(kp, e, f)
                                if wb:
                                    W("      %s += %s * %s;" % (acc_name(bp, kp, e, f), wb, t)); ops_bra.n += 2
                                else:
                                    W("      %s += %s;" % (acc_name(bp, kp, e, f), t)); ops_bra.n += 1
        W("    }")
        if hrr == "acc":
            (bp, kp), = combos
            return {(e, f): acc_name(bp, kp, e, f) for le in e_by_bp[bp] for e in cart(le) for f in fl(kp)}
        if hrr == "cd_loop":
            (bp, kp), = combos
            have = {(e, f): acc_name(bp, kp, e, f) for le in e_by_bp[bp] for e in cart(le) for f in fl(kp)}
            return self.emit_hrr_cd_loop(W, bp, kp, lambda e, f: have[(e, f)] if (e, f) in have else ext(e, f),
                                         ops_ctr, gid, sink)
        # ---- HRR + spherical per part combination ----
        elem = {}
        cA, cB, cC, cD = comps(self.A), comps(self.B), comps(self.C), comps(self.D)
        if sink is not None:
            for (bp, kp) in combos:
                self.emit_hrr_ket_first(W, bp, kp, acc_name, ops_ctr, gid, sink, elem)
            return elem
        for (bp, kp) in combos:
            la, lb, _ = self.erange(bp)
            lc, ld, fr = self.frange(kp)
            tag = "%s%d%d%d%d" % (gid, bp[0], bp[1], kp[0], kp[1])
            fl = [f for lf in fr for f in cart(lf)]
            hm = {}

            def hb(a_, b_, f_, bp=bp, kp=kp, hm=hm, tag=tag):
                key = (a_, b_, f_)
                if key in hm:
                    return hm[key]
                if sum(b_) == 0:
                    hm[key] = acc_name(bp, kp, a_, f_)
                    return hm[key]
                i = next(k for k in range(3) if b_[k] > 0)
                x1 = hb(add1(a_, i), sub1(b_, i), f_)
                x2 = hb(a_, sub1(b_, i), f_)
                nm = "hb%s_%s_%s_%s" % (tag, tname(a_), tname(b_), tname(f_))
                W("    const R %s = %s + AB%s * %s;" % (nm, x1, AX[i], x2))
                ops_ctr.n += 2
                hm[key] = nm
                return nm

            braname = {}
            for ia, ca in enumerate(cA):
                if ca[0] != bp[0]:
                    continue
                for ib, cb in enumerate(cB):
                    if cb[0] != bp[1]:
                        continue
                    for f_ in fl:
                        ta = [(ca[2], "1")] if ca[1] == "c" else SPH[ca[2]]
                        tb = [(cb[2], "1")] if cb[1] == "c" else SPH[cb[2]]
                        terms = [(hb(x, y, f_), fx, fy) for (x, fx) in ta for (y, fy) in tb]
                        if len(terms) == 1 and terms[0][1] == "1" and terms[0][2] == "1":
                            braname[(ia, ib, f_)] = terms[0][0]
                        else:
                            nm = "bs%s_%d_%d_%s" % (tag, ia, ib, tname(f_))
                            ex = " + ".join("(R)(%s*%s) * %s" % (fx, fy, t) if (fx, fy) != ("1", "1") else t
                                            for (t, fx, fy) in terms)
                            W("    const R %s = %s;" % (nm, ex))
                            ops_ctr.n += 2 * len(terms) - 1
                            braname[(ia, ib, f_)] = nm
            for ia, ca in enumerate(cA):
                if ca[0] != bp[0]:
                    continue
                for ib, cb in enumerate(cB):
                    if cb[0] != bp[1]:
                        continue
                    km = {}

                    def hk(c_, d_, ia=ia, ib=ib, km=km, tag=tag, braname=braname):
                        key = (c_, d_)
                        if key in km:
                            return km[key]
                        if sum(d_) == 0:
                            km[key] = braname[(ia, ib, c_)]
                            return km[key]
                        i = next(k for k in range(3) if d_[k] > 0)
                        x1 = hk(add1(c_, i), sub1(d_, i))
                        x2 = hk(c_, sub1(d_, i))
                        nm = "hk%s_%d_%d_%s_%s" % (tag, ia, ib, tname(c_), tname(d_))
                        W("    const R %s = %s + CD%s * %s;" % (nm, x1, AX[i], x2))
                        ops_ctr.n += 2
                        km[key] = nm
                        return nm

                    for ic, cc in enumerate(cC):
                        if cc[0] != kp[0]:
                            continue
                        for id_, cd in enumerate(cD):
                            if cd[0] != kp[1]:
                                continue
                            tc_ = [(cc[2], "1")] if cc[1] == "c" else SPH[cc[2]]
                            td_ = [(cd[2], "1")] if cd[1] == "c" else SPH[cd[2]]
                            terms = [(hk(x, y), fx, fy) for (x, fx) in tc_ for (y, fy) in td_]
                            nm = "e_%d_%d_%d_%d" % (ia, ib, ic, id_)
                            if len(terms) == 1 and terms[0][1] == "1" and terms[0][2] == "1":
                                W("    const R %s = %s;" % (nm, terms[0][0]))
                            else:
                                ex = " + ".join("(R)(%s*%s) * %s" % (fx, fy, t) if (fx, fy) != ("1", "1") else t
                                                for (t, fx, fy) in terms)
                                W("    const R %s = %s;" % (nm, ex))
                                ops_ctr.n += 2 * len(terms) - 1
                            elem[(ia, ib, ic, id_)] = nm
        return elem

    def emit_hrr_ket_first(self, W, bp, kp, acc_name, ops_ctr, gid, sink, elem):
        """HRR on the ket for every bra index e (with the ket's spherical
        transform), then per ket component the HRR on the bra, its spherical
        transform, and the element handed to `sink`:
          (e0|c,d+1_i) = (e0|c+1_i,d) + CD_i (e0|c,d),  base (e0|f0) = contracted VRR;
          (a,b+1_i|cd) = (a+1_i,b|cd) + AB_i (a,b|cd),  base (e0|cd)."""
        cA, cB, cC, cD = comps(self.A), comps(self.B), comps(self.C), comps(self.D)
        la, lb, er = self.erange(bp)
        tag = "%s%d%d%d%d" % (gid, bp[0], bp[1], kp[0], kp[1])
        el = [e for le in er for e in cart(le)]
        kets = [(ic, cc, id_, cd) for ic, cc in enumerate(cC) if cc[0] == kp[0]
                for id_, cd in enumerate(cD) if cd[0] == kp[1]]
        ketname = {}
        for e_ in el:
            km = {}

            def hk(c_, d_, e_=e_, km=km):
                key = (c_, d_)
                if key in km:
                    return km[key]
                if sum(d_) == 0:
                    km[key] = acc_name(bp, kp, e_, c_)
                    return km[key]
                i = next(k for k in range(3) if d_[k] > 0)
                x1 = hk(add1(c_, i), sub1(d_, i))
                x2 = hk(c_, sub1(d_, i))
                nm = "hk%s_%s_%s_%s" % (tag, tname(e_), tname(c_), tname(d_))
                W("    const R %s = %s + CD%s * %s;" % (nm, x1, AX[i], x2))
                ops_ctr.n += 2
                km[key] = nm
                return nm

            for (ic, cc, id_, cd) in kets:
                tc_ = [(cc[2], "1")] if cc[1] == "c" else SPH[cc[2]]
                td_ = [(cd[2], "1")] if cd[1] == "c" else SPH[cd[2]]
                terms = [(hk(x, y), fx, fy) for (x, fx) in tc_ for (y, fy) in td_]
                if len(terms) == 1 and terms[0][1] == "1" and terms[0][2] == "1":
                    ketname[(e_, ic, id_)] = terms[0][0]
                else:
                    nm = "ks%s_%s_%d_%d" % (tag, tname(e_), ic, id_)
                    ex = " + ".join("(R)(%s*%s) * %s" % (fx, fy, t) if (fx, fy) != ("1", "1") else t
                                    for (t, fx, fy) in terms)
                    W("    const R %s = %s;" % (nm, ex))
                    ops_ctr.n += 2 * len(terms) - 1
                    ketname[(e_, ic, id_)] = nm
        for (ic, cc, id_, cd) in kets:
            hm = {}

            def hb(a_, b_, ic=ic, id_=id_, hm=hm):
                key = (a_, b_)
                if key in hm:
                    return hm[key]
                if sum(b_) == 0:
                    hm[key] = ketname[(a_, ic, id_)]
                    return hm[key]
                i = next(k for k in range(3) if b_[k] > 0)
                x1 = hb(add1(a_, i), sub1(b_, i))
                x2 = hb(a_, sub1(b_, i))
                nm = "hb%s_%d_%d_%s_%s" % (tag, ic, id_, tname(a_), tname(b_))
                W("    const R %s = %s + AB%s * %s;" % (nm, x1, AX[i], x2))
                ops_ctr.n += 2
                hm[key] = nm
                return nm

            for ia, ca in enumerate(cA):
                if ca[0] != bp[0]:
                    continue
                for ib, cb in enumerate(cB):
                    if cb[0] != bp[1]:
                        continue
                    ta = [(ca[2], "1")] if ca[1] == "c" else SPH[ca[2]]
                    tb = [(cb[2], "1")] if cb[1] == "c" else SPH[cb[2]]
                    terms = [(hb(x, y), fx, fy) for (x, fx) in ta for (y, fy) in tb]
                    nm = "e_%d_%d_%d_%d" % (ia, ib, ic, id_)
                    if len(terms) == 1 and terms[0][1] == "1" and terms[0][2] == "1":
                        W("    const R %s = %s;" % (nm, terms[0][0]))
                    else:
                        ex = " + ".join("(R)(%s*%s) * %s" % (fx, fy, t) if (fx, fy) != ("1", "1") else t
                                        for (t, fx, fy) in terms)
                        W("    const R %s = %s;" % (nm, ex))
                        ops_ctr.n += 2 * len(terms) - 1
                    elem[(ia, ib, ic, id_)] = nm
                    sink(ia, ib, ic, id_, nm)

    def emit_hrr_cd_loop(self, W, bp, kp, val, ops_ctr, gid, sink):
        """HRR one ket output (ic, id) at a time, so that besides the
        contracted [e0|f0] only the current output's (e|cd) values are live:
          (e0|c,d+1_i) = (e0|c+1_i,d) + CD_i (e0|c,d)   for every bra index e,
          (a,b+1_i|cd) = (a+1_i,b|cd) + AB_i (a,b|cd),
        with the spherical transforms; val(e, f) is the expression of [e0|f0]
        (an accumulator, or a load of what another kernel stored).  Ket
        intermediates are recomputed per ket output (none are shared between
        outputs for a p or s D shell)."""
        cA, cB, cC, cD = comps(self.A), comps(self.B), comps(self.C), comps(self.D)
        la, lb, er = self.erange(bp)
        tag = "%s%d%d%d%d" % (gid, bp[0], bp[1], kp[0], kp[1])
        el = [e for le in er for e in cart(le)]
        kets = [(ic, cc, id_, cd) for ic, cc in enumerate(cC) if cc[0] == kp[0]
                for id_, cd in enumerate(cD) if cd[0] == kp[1]]
        elem = {}
        for (ic, cc, id_, cd) in kets:
            W("    {")
            tc_ = [(cc[2], "1")] if cc[1] == "c" else SPH[cc[2]]
            td_ = [(cd[2], "1")] if cd[1] == "c" else SPH[cd[2]]
            base = {}    # [e0|f0] loaded or named once per ket output

            def b(e_, f_):
                key = (e_, f_)
                if key not in base:
                    nm = "cb%s_%s_%s" % (tag, tname(e_), tname(f_))
                    W("      const R %s = %s;" % (nm, val(e_, f_)))
                    base[key] = nm
                return base[key]

            ketname = {}
            for e_ in el:
                km = {}

                def hk(c_, d_, e_=e_, km=km):
                    key = (c_, d_)
                    if key in km:
                        return km[key]
                    if sum(d_) == 0:
                        km[key] = b(e_, c_)
                        return km[key]
                    i = next(k for k in range(3) if d_[k] > 0)
                    x1 = hk(add1(c_, i), sub1(d_, i))
                    x2 = hk(c_, sub1(d_, i))
                    nm = "hk%s_%s_%s_%s" % (tag, tname(e_), tname(c_), tname(d_))
                    W("      const R %s = %s + CD%s * %s;" % (nm, x1, AX[i], x2))
                    ops_ctr.n += 2
                    km[key] = nm
                    return nm

                terms = [(hk(x, y), fx, fy) for (x, fx) in tc_ for (y, fy) in td_]
                if len(terms) == 1 and terms[0][1] == "1" and terms[0][2] == "1":
                    ketname[e_] = terms[0][0]
                else:
                    nm = "ks%s_%s" % (tag, tname(e_))
                    ex = " + ".join("(R)(%s*%s) * %s" % (fx, fy, t) if (fx, fy) != ("1", "1") else t
                                    for (t, fx, fy) in terms)
                    W("      const R %s = %s;" % (nm, ex))
                    ops_ctr.n += 2 * len(terms) - 1
                    ketname[e_] = nm
            hm = {}

            def hb(a_, b_, hm=hm, ketname=ketname):
                key = (a_, b_)
                if key in hm:
                    return hm[key]
                if sum(b_) == 0:
                    hm[key] = ketname[a_]
                    return hm[key]
                i = next(k for k in range(3) if b_[k] > 0)
                x1 = hb(add1(a_, i), sub1(b_, i))
                x2 = hb(a_, sub1(b_, i))
                nm = "hb%s_%s_%s" % (tag, tname(a_), tname(b_))
                W("      const R %s = %s + AB%s * %s;" % (nm, x1, AX[i], x2))
                ops_ctr.n += 2
                hm[key] = nm
                return nm

            for ia, ca in enumerate(cA):
                if ca[0] != bp[0]:
                    continue
                for ib, cb in enumerate(cB):
                    if cb[0] != bp[1]:
                        continue
                    ta = [(ca[2], "1")] if ca[1] == "c" else SPH[ca[2]]
                    tb = [(cb[2], "1")] if cb[1] == "c" else SPH[cb[2]]
                    terms = [(hb(x, y), fx, fy) for (x, fx) in ta for (y, fy) in tb]
                    nm = "e_%d_%d_%d_%d" % (ia, ib, ic, id_)
                    if len(terms) == 1 and terms[0][1] == "1" and terms[0][2] == "1":
                        W("      const R %s = %s;" % (nm, terms[0][0]))
                    else:
                        ex = " + ".join("(R)(%s*%s) * %s" % (fx, fy, t) if (fx, fy) != ("1", "1") else t
                                        for (t, fx, fy) in terms)
                        W("      const R %s = %s;" % (nm, ex))
                        ops_ctr.n += 2 * len(terms) - 1
                    elem[(ia, ib, ic, id_)] = nm
                    sink(ia, ib, ic, id_, nm)
            W("    }")
        return elem

    def ref_index(self, ia, ib, ic, id_):
        n_ref = [NCOMP[x] for x in self.cls]
        i1, i2 = (ib, ia) if self.swap_bra else (ia, ib)
        i3, i4 = (id_, ic) if self.swap_ket else (ic, id_)
        return i1 + n_ref[0] * (i2 + n_ref[1] * (i3 + n_ref[2] * i4))

    def dims(self):
        return [NCOMP[x] for x in (self.A, self.B, self.C, self.D)]

    def group_plan(self, split, fp32=False):
        """Part combinations per kernel: all of them (fused classes); for the
        block kernels, consecutive combinations of one bra part until their
        contracted accumulators exceed the budget (FP64_GROUP_ACC, or
        FP32_GROUP_ACC for FP32, whose values take half the registers): fewer
        kernels, so fewer passes over the primitive pairs and Boys
        evaluations, while a heavy combination keeps a kernel, and the
        register budget, of its own."""
        combos = [(bp, kp) for bp in self.bra_parts for kp in self.ket_parts]
        if not split:
            return [combos]
        b64, b32 = GROUP_BUDGETS.get(self.cls, (FP64_GROUP_ACC, FP32_GROUP_ACC))
        budget = b32 if fp32 else b64

        def nc(ls):
            return sum((l + 1) * (l + 2) // 2 for l in ls)
        groups = []
        for bp in self.bra_parts:
            cur, cost = [], 0
            for kp in self.ket_parts:
                c = nc(self.erange(bp)[2]) * nc(self.frange(kp)[2])
                if cur and cost + c > budget:
                    groups.append(cur)
                    cur, cost = [], 0
                cur.append((bp, kp))
                cost += c
            if cur:
                groups.append(cur)
        return groups

    def split_parts(self, combos, fp32=False):
        """A group that is one part combination whose contracted accumulators
        exceed SPLIT_ACC (FP64) or 2 x SPLIT_ACC (FP32): its ket components
        f in near-equal parts, each part's [e0|f0] held by one kernel (the
        pre-kernels store theirs, the last kernel runs the HRR); else None."""
        if len(combos) != 1 or not SPLIT_CLASSES.get(self.cls, (False, False))[1 if fp32 else 0]:
            return None
        (bp, kp), = combos
        ne = sum((l + 1) * (l + 2) // 2 for l in self.erange(bp)[2])
        fl = [f for lf in self.frange(kp)[2] for f in cart(lf)]
        budget = SPLIT_ACC * (2 if fp32 else 1)
        if ne * len(fl) <= budget:
            return None
        k = -(-ne * len(fl) // budget)          # parts
        size = -(-len(fl) // k)
        return [fl[i:i + size] for i in range(0, len(fl), size)]

    def group_M(self, combos):
        return max(self.combo_M(c) for c in combos)

    def kernel_fused(self, W, ops_prim, ops_bra, ops_ctr):
        """One thread per quartet; integrals, block test and J/K digestion in
        registers (light classes).  With a.lds_rows the workgroup walks a
        contiguous range of the bra-contiguous list, 256 quartets per step,
        and keeps the rows of X of one bra pair (the one owning most of the
        step's quartets, pick_resident) in LDS after the Boys table: J_ab (summed over the wave
        first), K_ac, K_ad, K_bc and K_bd of that bra's quartets add there and
        go to X with row-coalesced atomics when the bra changes and at the
        end; J_cd and other quartets add to X directly.  Without it (diagonal
        builds, rows larger than the LDS allows) every sum is an atomic to X."""
        cls = self.cls
        M = self.M
        nA, nB, nC, nD = self.dims()
        oa = "b" if self.swap_bra else "a"
        ob = "a" if self.swap_bra else "b"
        W("template <typename R, bool JONLY>")
        W("__global__ void __launch_bounds__(HFX_NT) eri_%s(const ClassArgs a) {" % cls)
        W("  extern __shared__ char hfx_smem[];")
        W("  R* btab = reinterpret_cast<R*>(hfx_smem);")
        W("  constexpr int NB_COL = %d + BoysTerms<R>::K;" % M)
        W("  stage_boys<R>(btab, a, NB_COL);")
        W("  const int nq = *a.count;")
        W("  const int nbf = a.nb;")
        W("  const bool rows = a.lds_rows != 0;")
        W("  double* srow = reinterpret_cast<double*>(hfx_smem + a.rows_offset);   // [%d][nb]: rows A then B" % (nA + nB))
        W("  const int lane = threadIdx.x & (HFX_WAVE - 1);")
        W("  double* X = a.X;")
        W("  const int nsteps = (nq + blockDim.x - 1) / blockDim.x;")
        W("  int s0, s1;")
        W("  if (rows) {   // contiguous range of steps per workgroup")
        W("    const int per = (nsteps + gridDim.x - 1) / gridDim.x;")
        W("    s0 = blockIdx.x * per; s1 = min(nsteps, s0 + per);")
        W("    for (int e = threadIdx.x; e < %d * nbf; e += blockDim.x) srow[e] = 0.0;" % (nA + nB))
        W("  } else {")
        W("    s0 = blockIdx.x; s1 = nsteps;")
        W("  }")
        W("  __syncthreads();   // rows zeroed before any wave adds to them")
        W("  const int sstride = rows ? 1 : gridDim.x;")
        W("  int curb = -1, rA = 0, rB = 0;   // resident bra pair and its rows")
        W("  double* rX = X;                  // its molecule's block of X")
        W("  auto flush = [&]() {")
        W("    __syncthreads();")
        W("    for (int e = threadIdx.x; e < %d * nbf; e += blockDim.x) {" % (nA + nB))
        W("      const double v = srow[e];")
        W("      if (v != 0.0) {")
        W("        const int r = e / nbf, col = e - r * nbf;")
        W("        const int row = r < %d ? rA + r : rB + (r - %d);" % (nA, nA))
        W("        atomicAdd(&rX[(size_t)row * nbf + col], v);")
        W("        srow[e] = 0.0;")
        W("      }")
        W("    }")
        W("    __syncthreads();")
        W("  };")
        W("  for (int st = s0; st < s1; st += sstride) {")
        W("    if (rows) {")
        W("      const int tb = pick_resident(a.list, st * (int)blockDim.x, min((st + 1) * (int)blockDim.x, nq), curb);")
        W("      if (tb != curb) {")
        W("        if (curb >= 0) flush();")
        W("        curb = tb;")
        W("        rA = a.pairs[tb].ao%s; rB = a.pairs[tb].ao%s;" % (oa, ob))
        W("        rX = X + a.pairs[tb].mbase;")
        W("      }")
        W("    }")
        W("    const int it = st * blockDim.x + (threadIdx.x - lane);   // first quartet of the wave")
        W("    if (it >= nq) continue;")
        W("    const bool valid = it + lane < nq;")
        W("    const int iq = valid ? it + lane : nq - 1;")
        W("    const int2 bk = a.list[iq];")
        self.emit_quartet_setup(W)
        elem = {}
        for gi, combos in enumerate(self.group_plan(False)):
            elem.update(self.emit_group(W, combos, ops_prim, ops_bra, ops_ctr, "g%d" % gi))
        assert len(elem) == nA * nB * nC * nD

        def pblk(name, ao1, n1, ao2, n2, konly):
            for i in range(n1):
                for j in range(n2):
                    ld = "Pq[(size_t)(%s + %d) * nbf + %s + %d]" % (ao1, i, ao2, j)
                    if konly:
                        ld = "JONLY ? 0.0 : " + ld
                    W("    const double %s_%d_%d = %s;" % (name, i, j, ld))

        W("    const double* Pq = a.P + mq;   // the quartet's molecule's block of P and X")
        W("    double* Xq = X + mq;")
        pblk("Pab", "aoA", nA, "aoB", nB, False)
        pblk("Pcd", "aoC", nC, "aoD", nD, False)
        pblk("Pac", "aoA", nA, "aoC", nC, True)
        pblk("Pad", "aoA", nA, "aoD", nD, True)
        pblk("Pbc", "aoB", nB, "aoC", nC, True)
        pblk("Pbd", "aoB", nB, "aoD", nD, True)

        def decl(name, n1, n2):
            for i in range(n1):
                for j in range(n2):
                    W("    double %s_%d_%d = 0;" % (name, i, j))

        decl("Jab", nA, nB); decl("Jcd", nC, nD)
        decl("Kac", nA, nC); decl("Kad", nA, nD); decl("Kbc", nB, nC); decl("Kbd", nB, nD)
        W("    R vmax = 0;")
        W("    const R ecut = a.diag ? (R)HFX_INTTHRESH : (R)0;")
        W("    const bool dump = a.blocks != nullptr && valid;")
        for (ia, ib, ic, id_), nm in sorted(elem.items()):
            W("    {")
            W("      R v = %s;" % nm)
            W("      vmax = fmax(vmax, fabs(v));")
            W("      if (fabs(v) <= ecut) v = 0;")
            W("      if (dump) a.blocks[(size_t)iq * 625 + %d] = (double)v;" % self.ref_index(ia, ib, ic, id_))
            W("      const double dv = (double)v;")
            W("      Jab_%d_%d = fma(dv, Pcd_%d_%d, Jab_%d_%d);" % (ia, ib, ic, id_, ia, ib))
            W("      Jcd_%d_%d = fma(dv, Pab_%d_%d, Jcd_%d_%d);" % (ic, id_, ia, ib, ic, id_))
            W("      if constexpr (!JONLY) {")
            W("        Kac_%d_%d = fma(dv, Pbd_%d_%d, Kac_%d_%d);" % (ia, ic, ib, id_, ia, ic))
            W("        Kad_%d_%d = fma(dv, Pbc_%d_%d, Kad_%d_%d);" % (ia, id_, ib, ic, ia, id_))
            W("        Kbc_%d_%d = fma(dv, Pad_%d_%d, Kbc_%d_%d);" % (ib, ic, ia, id_, ib, ic))
            W("        Kbd_%d_%d = fma(dv, Pac_%d_%d, Kbd_%d_%d);" % (ib, id_, ia, ic, ib, id_))
            W("      }")
            W("    }")
            ops_ctr.n += 12 + 1
        W("    if (a.schw_out && valid) a.schw_out[bk.x] = sqrt((double)vmax);")
        W("    const double thq = a.thr_mol ? a.thr_mol[pb.mol] : a.thr;")
        W("    const bool act = valid && (a.diag || (double)vmax >= thq);")
        W("    if (__ballot(act) == 0) continue;")
        W("    const bool inl = rows && act && bk.x == curb;   // rows A, B resident in LDS")
        W("    const double w = act ? ((shA == shB) ? 0.5 : 1.0) * ((shC == shD) ? 0.5 : 1.0) * (a.diag ? 0.5 : 1.0) : 0.0;")
        W("    const double wj = 2.0 * w, wk = -0.5 * w;")
        for i in range(nA):
            for j in range(nB):
                W("    {")
                W("      const double jab = wj * Jab_%d_%d;" % (i, j))
                W("      if (rows) {")
                W("        double r = inl ? jab : 0.0;")
                W("        r = wave_sum(r);")
                W("        if (lane == 0 && r != 0.0) atomicAdd(&srow[%d * nbf + rB + %d], r);" % (i, j))
                W("      }")
                W("      if (act && !inl) atomicAdd(&Xq[(size_t)(aoA + %d) * nbf + aoB + %d], jab);" % (i, j))
                W("    }")
        # J_cd into the ket pair's contiguous pair-buffer slot (the pair's own shell order)
        W("    if (act) {")
        W("      double* pbq = a.pairbuf + (size_t)bk.y * HFX_PAIRBUF_STRIDE;")
        for i in range(nC):
            for j in range(nD):
                pe = j * nC + i if self.swap_ket else i * nD + j
                W("      atomicAdd(&pbq[%d], wj * Jcd_%d_%d);" % (pe, i, j))
        W("    }")

        def krow(name, ao1, n1, row0, ao2, n2):
            W("      if (inl) {")
            for i in range(n1):
                for j in range(n2):
                    W("        atomicAdd(&srow[%d * nbf + %s + %d], wk * %s_%d_%d);" % (row0 + i, ao2, j, name, i, j))
            W("      } else if (act) {")
            for i in range(n1):
                for j in range(n2):
                    W("        atomicAdd(&Xq[(size_t)(%s + %d) * nbf + %s + %d], wk * %s_%d_%d);" % (ao1, i, ao2, j, name, i, j))
            W("      }")

        W("    if constexpr (!JONLY) {")
        krow("Kac", "aoA", nA, 0, "aoC", nC)
        krow("Kad", "aoA", nA, 0, "aoD", nD)
        krow("Kbc", "aoB", nB, nA, "aoC", nC)
        krow("Kbd", "aoB", nB, nA, "aoD", nD)
        W("    }")
        W("  }")
        W("  if (rows && curb >= 0) flush();")
        W("}")

    def combo_M(self, combo):
        bp, kp = combo
        return max(self.erange(bp)[2]) + max(self.frange(kp)[2])

    def kernel_block(self, W, ops_prim, ops_bra, ops_ctr, fp32=False, plan=None, group_ops=None):
        """One kernel per group of L-shell part combinations (group_plan):
        one thread per quartet of the chunk [a.q0, a.q1), its sub-block stored
        to a.eblk (quartet-major, oriented layout); digestion follows in
        dig_<class>.  FP64 kernels are eri_<class>_c<k>, FP32 ones
        eri_<class>_f<k>."""
        cls = self.cls
        nA, nB, nC, nD = self.dims()
        n = 0
        kname = "f" if fp32 else "c"
        goff = 0   # contracted values stored by the pre-kernels of the class, per quartet, so far

        def head(name, M):
            W("template <typename R>")
            W("__global__ void __launch_bounds__(HFX_NT) %s(const ClassArgs a) {" % name)
            W("  extern __shared__ char hfx_smem[];")
            W("  if (a.q0 >= *a.count) return;   // a chunk past the list's end (chunks sized by capacity)")
            W("  R* btab = reinterpret_cast<R*>(hfx_smem);")
            W("  constexpr int NB_COL = %d + BoysTerms<R>::K;" % M)
            W("  stage_boys<R>(btab, a, NB_COL);")
            W("  const int qend = min(a.q1, *a.count);")
            W("  for (int iq = a.q0 + blockIdx.x * blockDim.x + threadIdx.x; iq < qend; iq += gridDim.x * blockDim.x) {")
            W("    const int2 bk = a.list[iq];")
            W("    const size_t slot = (size_t)(iq - a.q0);")
            W("    R* cb = reinterpret_cast<R*>(a.cbuf); (void)cb;   // contracted values of split combinations")

        def tail(name):
            W("  }")
            W("}")
            W("template __global__ void %s<%s>(const ClassArgs);" % (name, "float" if fp32 else "double"))

        for gi, combos in enumerate(plan if plan is not None else self.group_plan(True, fp32)):
            M = self.group_M(combos)
            parts = self.split_parts(combos, fp32)
            kern = "eri_%s_%s%d" % (cls, kname, gi)
            ext = None
            fsel = None
            if parts is not None:
                # pre-kernels: parts[:-1] of the ket components, contracted and stored
                (bp, kp), = combos
                el = [e for le in self.erange(bp)[2] for e in cart(le)]
                stored = [f for part in parts[:-1] for f in part]
                idx = {}
                for e in el:
                    for f in stored:
                        idx[(e, f)] = goff + len(idx)
                for pi, part in enumerate(parts[:-1]):
                    pname = "%sv%d" % (kern, pi)
                    head(pname, M)
                    self.emit_quartet_setup(W)
                    acc = self.emit_group(W, combos, ops_prim, ops_bra, ops_ctr, "g%dv%d" % (gi, pi),
                                          fsel=set(part), hrr="acc")
                    for (e, f), nm in acc.items():
                        W("    cb[(size_t)%d * a.vstride + slot] = %s;" % (idx[(e, f)], nm))
                    tail(pname)
                goff += len(idx)
                fsel = set(parts[-1])
                ext = lambda e, f, idx=idx: "cb[(size_t)%d * a.vstride + slot]" % idx[(e, f)]
            head(kern, M)
            W("    R* out = reinterpret_cast<R*>(a.eblk) + (slot >> 6) * %d + (slot & 63);" % (nA * nB * nC * nD * 64))
            self.emit_quartet_setup(W)
            # raw values and their partial max|e| (the block test); the
            # diagonal builds' element cut runs on the stored blocks
            # (k_cut_small)
            W("    R vm = 0;   // in the block's precision: FP32 maxima convert exactly")

            def store(ia, ib, ic, id_, nm):
                W("    { const R o = %s; vm = fmax(vm, fabs(o)); out[%d] = o; }" % (
                    nm, (((ia * nB + ib) * nC + ic) * nD + id_) * 64))

            if parts is None:
                el = self.emit_group(W, combos, ops_prim, ops_bra, ops_ctr, "g%d" % gi, sink=store)
            else:
                el = self.emit_group(W, combos, ops_prim, ops_bra, ops_ctr, "g%d" % gi, sink=store, fsel=fsel,
                                     hrr="cd_loop", ext=ext)
            if group_ops is not None:
                group_ops.append((ops_prim.n, ops_bra.n, ops_ctr.n))
            W("    a.vmax[(size_t)%d * a.vstride + slot] = (double)vm;" % gi)
            n += len(el)
            tail(kern)
        assert n == nA * nB * nC * nD
        return goff

    def kernel_digest(self, W, ops_ctr):
        """Digestion of the stored blocks of one block class: one thread per
        quartet (tile-major blocks, so every element load is one coalesced
        wave access).  The block test (max|e| < thr skips the quartet,
        fock_build_non_diag.f90:69) uses the partial maxima the combination
        kernels stored.  Two passes over the block (three, see below), each flushing an
        accumulator as soon as it is complete:
          J:  X_ab += 2w sum_cd e P_cd (per (a,b)),  X_cd += 2w sum_ab e P_ab;
          K:  X_ac -= w/2 sum_bd e P_bd, X_ad -= w/2 sum_bc e P_bc (per a),
              X_bc -= w/2 sum_ad e P_ad, X_bd -= w/2 sum_ac e P_ac (all b, after the last a);
        or, for the classes in KPASS_TWO, K in two passes (K1 per a: X_ac,
        X_ad; K2 per b: X_bc, X_bd), which hold fewer registers.
        Rows of A and B: with d.lds_rows the workgroup walks a contiguous
        range of the bra-contiguous list 256 quartets at a time and keeps the
        rows of X of one bra pair (the one owning most of the step's
        quartets, pick_resident) in LDS; quartets of that bra add there (X_ab
        summed over the wave first), the rows go to X with row-coalesced
        atomics when the bra changes and at the end.  Other quartets add to X
        directly; X_cd goes to the ket pair's slot of the pair buffer.
        A compiler barrier after each slice of loads keeps at most one slice
        of the block in registers."""
        cls = self.cls
        nA, nB, nC, nD = self.dims()
        ne = nA * nB * nC * nD
        ncombo = len(self.group_plan(True))
        oa = "b" if self.swap_bra else "a"
        ob = "a" if self.swap_bra else "b"
        oc = "b" if self.swap_ket else "a"
        od = "a" if self.swap_ket else "b"

        def lin(ia, ib, ic, id_):
            return (((ia * nB + ib) * nC + ic) * nD + id_) * 64

        def pload(name, ao1, n1, ao2, n2):
            for i in range(n1):
                for j in range(n2):
                    W("      const T %s_%d_%d = PQ[(size_t)(%s + %d) * nbf + %s + %d];" % (name, i, j, ao1, i, ao2, j))

        def zero(name, n):
            W("      " + " ".join("T %s_%d = 0;" % (name, i) for i in range(n)))

        def pin(name, n):
            W("      " + " ".join("HFX_PIN(%s_%d);" % (name, i) for i in range(n)) + " HFX_LOAD_FENCE();")

        def flush_row(name, n, ao1, i, row, ao2, wname):
            """accumulator row `name` (n columns from ao2) of X row ao1 + i;
            LDS row `row` when the quartet's bra is the resident one"""
            W("      if (inl) {")
            for j in range(n):
                W("        atomicAdd(&srow[%d * nbf + %s + %d], %s * (double)%s_%d);" % (row, ao2, j, wname, name, j))
            W("      } else if (act) {")
            for j in range(n):
                W("        atomicAdd(&XQ[(size_t)(%s + %d) * nbf + %s + %d], %s * (double)%s_%d);" % (ao1, i, ao2, j, wname, name, j))
            W("      }")

        # MULTI: molecule-aware form for batch contexts (P and X block and
        # threshold per quartet's molecule); the single-molecule form keeps
        # the uniform base pointers and threshold, which the register-bound
        # two-pass classes measured faster with ((dl|dl) 4.2 -> 2.8 ms)
        W("template <bool JONLY, typename T, bool MULTI>")
        W("__global__ void __launch_bounds__(HFX_NT) dig_%s(const DigestArgs d) {" % cls)
        W("  extern __shared__ char hfx_smem[];")
        W("  double* srow = reinterpret_cast<double*>(hfx_smem);   // [%d][nb]: rows A then rows B of X" % (nA + nB))
        W("  const int qend = min(d.q1, *d.count);")
        W("  if (qend <= d.q0) return;")
        W("  const int nbf = d.nb;")
        W("  const bool rows = d.lds_rows != 0;")
        W("  const int lane = threadIdx.x & (HFX_WAVE - 1);")
        W("  double* X = d.X;")
        W("  // contractions in the element type: FP64 blocks with P, FP32 blocks with its FP32 copy;")
        W("  // each quartet's sums go to X in FP64")
        W("  const T* PA;")
        W("  if constexpr (sizeof(T) == sizeof(float)) PA = reinterpret_cast<const T*>(d.P32);")
        W("  else PA = reinterpret_cast<const T*>(d.P);")
        W("  const int nsteps = (qend - d.q0 + blockDim.x - 1) / blockDim.x;")
        W("  int s0, s1;")
        W("  if (rows) {   // contiguous range of steps per workgroup")
        W("    const int per = (nsteps + gridDim.x - 1) / gridDim.x;")
        W("    s0 = blockIdx.x * per; s1 = min(nsteps, s0 + per);")
        W("    for (int e = threadIdx.x; e < %d * nbf; e += blockDim.x) srow[e] = 0.0;" % (nA + nB))
        W("  } else {")
        W("    s0 = blockIdx.x; s1 = nsteps;")
        W("  }")
        W("  __syncthreads();   // rows zeroed before any wave adds to them")
        W("  const int sstride = rows ? 1 : gridDim.x;")
        W("  int curb = -1, rA = 0, rB = 0;   // resident bra pair and its rows")
        W("  double* rX = X;                  // its molecule's block of X")
        W("  auto flush = [&]() {")
        W("    __syncthreads();")
        W("    for (int e = threadIdx.x; e < %d * nbf; e += blockDim.x) {" % (nA + nB))
        W("      const double v = srow[e];")
        W("      if (v != 0.0) {")
        W("        const int r = e / nbf, col = e - r * nbf;")
        W("        const int row = r < %d ? rA + r : rB + (r - %d);" % (nA, nA))
        W("        atomicAdd(&rX[(size_t)row * nbf + col], v);")
        W("        srow[e] = 0.0;")
        W("      }")
        W("    }")
        W("    __syncthreads();")
        W("  };")
        W("  for (int st = s0; st < s1; st += sstride) {")
        W("    const int it = d.q0 + st * blockDim.x + (threadIdx.x - lane);   // first quartet of the wave")
        W("    if (rows) {")
        W("      const int tb = pick_resident(d.list, d.q0 + st * (int)blockDim.x,")
        W("                                   min(d.q0 + (st + 1) * (int)blockDim.x, qend), curb);")
        W("      if (tb != curb) {")
        W("        if (curb >= 0) flush();")
        W("        curb = tb;")
        W("        rA = d.pairs[tb].ao%s; rB = d.pairs[tb].ao%s;" % (oa, ob))
        W("        if constexpr (MULTI) rX = X + d.pairs[tb].mbase;")
        W("      }")
        W("    }")
        W("    if (it >= qend) continue;")
        W("    const bool valid = it + lane < qend;")
        W("    const int iq = valid ? it + lane : qend - 1;")
        W("    const int slot = iq - d.q0;")
        W("    const int2 bk = d.list[iq];")
        W("    double vmax = d.vmax[slot];")
        W("    for (int c = 1; c < d.ncombo; ++c) vmax = fmax(vmax, d.vmax[(size_t)c * d.vstride + slot]);")
        W("    if (d.schw_out != nullptr && valid) d.schw_out[bk.x] = sqrt(vmax);")
        W("    double thq = d.thr;   // one molecule: its threshold is uniform (thr_mol[0] when given)")
        W("    if (d.thr_mol) thq = MULTI ? d.thr_mol[d.pairs[bk.x].mol] : d.thr_mol[0];")
        W("    const bool act = valid && (d.diag || vmax >= thq);")
        W("    const T* blk = reinterpret_cast<const T*>(d.eblk) + (size_t)(slot >> 6) * %d + (slot & 63);" % (ne * 64))
        W("    if (d.dump != nullptr && valid) {")
        W("      #pragma unroll 1")
        W("      for (int e = 0; e < %d; ++e) {" % ne)
        W("        const int id = e %% %d, ic = (e / %d) %% %d, ib = (e / %d) %% %d, ia = e / %d;" % (
            nD, nD, nC, nD * nC, nB, nD * nC * nB))
        i1, i2 = ("ib", "ia") if self.swap_bra else ("ia", "ib")
        i3, i4 = ("id", "ic") if self.swap_ket else ("ic", "id")
        m1, m2 = (nB, nA) if self.swap_bra else (nA, nB)
        m3 = nD if self.swap_ket else nC
        W("        d.dump[(size_t)iq * 625 + %s + %d * (%s + %d * (%s + %d * %s))] = (double)blk[(size_t)e * 64];" % (
            i1, m1, i2, m2, i3, m3, i4))
        W("      }")
        W("    }")
        W("    if (__ballot(act) == 0) continue;")
        W("    const bool inl = rows && act && bk.x == curb;   // rows A, B resident in LDS")
        W("    const PairRec& pb = d.pairs[bk.x];")
        W("    const PairRec& pk = d.pairs[bk.y];")
        W("    const int shA = pb.sh%s, shB = pb.sh%s, shC = pk.sh%s, shD = pk.sh%s;" % (oa, ob, oc, od))
        W("    const int aoA = pb.ao%s, aoB = pb.ao%s, aoC = pk.ao%s, aoD = pk.ao%s;" % (oa, ob, oc, od))
        W("    const T* PQ = MULTI ? PA + pb.mbase : PA;     // the quartet's molecule's block of P and X")
        W("    double* XQ = MULTI ? X + pb.mbase : X;")
        W("    const double w = act ? ((shA == shB) ? 0.5 : 1.0) * ((shC == shD) ? 0.5 : 1.0) * (d.diag ? 0.5 : 1.0) : 0.0;")
        W("    const double wj = 2.0 * w, wk = -0.5 * w;")
        ncd = nC * nD
        gs = 2 if ncd <= DIG_SLICE_PAIR_MAX else 1   # slices per fence (larger ones would need 2x registers)

        def jab_flush(ia, ib):
            W("      const double jd = wj * (double)jab;")
            W("      if (rows) {")
            W("        double r = inl ? jd : 0.0;")
            W("        r = wave_sum(r);")
            W("        if (lane == 0 && r != 0.0) atomicAdd(&srow[%d * nbf + rB + %d], r);" % (ia, ib))
            W("      }")
            W("      if (act && !inl) atomicAdd(&XQ[(size_t)(aoA + %d) * nbf + aoB + %d], jd);" % (ia, ib))

        # Pass J, then K below; each pass re-derives the block pointer
        # so that element addresses are not shared, and kept live, across
        # passes.  (Measured on C60 and dropped: J and K1 in one pass, and all
        # of J and K in one pass with per-(a, b) partial K sums in the LDS
        # rows -- fewer block reads, but 256 registers and one wave per SIMD,
        # slower overall.)
        W("    {")
        pload("Pcd", "aoC", nC, "aoD", nD)
        zero("Jcd", ncd)
        W("      const T* bp = blk; HFX_PIN_GPTR(bp);")
        for ia in range(nA):
            W("      {")
            for ib in range(nB):
                W("      {")
                W("      const T pab = PQ[(size_t)(aoA + %d) * nbf + aoB + %d];" % (ia, ib))
                W("      T jab = 0;")
                for ic in range(nC):
                    for id_ in range(nD):
                        W("      { const T v = bp[%d]; jab = fma(v, Pcd_%d_%d, jab); Jcd_%d = fma(v, pab, Jcd_%d); }" % (
                            lin(ia, ib, ic, id_), ic, id_, ic * nD + id_, ic * nD + id_))
                jab_flush(ia, ib)
                if (ia * nB + ib + 1) % gs == 0 or (ia == nA - 1 and ib == nB - 1):
                    pin("Jcd", ncd)
                W("      }")
            W("      }")
        # X_cd through LDS: TG quartets of the wave at a time, lanes over
        # consecutive elements, so one atomic instruction covers the
        # contiguous pair-buffer slots of a few quartets instead of 64
        # unrelated addresses
        W("      {")
        W("        constexpr int TG = HFX_WAVE < 16 ? HFX_WAVE : 16;")
        W("        double* tb = reinterpret_cast<double*>(hfx_smem + d.tb_offset) + (threadIdx.x / HFX_WAVE) * (TG * %d);" % ncd)
        W("        const double sc = act ? wj : 0.0;")
        W("        for (int r = 0; r < HFX_WAVE / TG; ++r) {")
        W("          if (lane / TG == r) {")
        W("            double* t = tb + (lane - r * TG) * %d;" % ncd)
        W("            " + " ".join("t[%d] = sc * (double)Jcd_%d;" % (k, k) for k in range(ncd)))
        W("          }")
        W("          wave_sync();")
        W("          for (int f0 = 0; f0 < TG * %d; f0 += HFX_WAVE) {" % ncd)
        W("            const int f = f0 + lane;")
        W("            const bool ok = f < TG * %d;" % ncd)
        W("            const int ql = ok ? f / %d : 0, e = ok ? f - ql * %d : 0;" % (ncd, ncd))
        W("            const int src = r * TG + ql;")
        # into the ket pair's slot of the pair buffer (contiguous, in the pair's own
        # shell order), added to X once per build
        W("            const int pq = __shfl(bk.y, src);")
        W("            const int pe = %s;" % ("(e %% %d) * %d + e / %d" % (nD, nC, nD) if self.swap_ket else "e"))
        W("            const double v = ok ? tb[f] : 0.0;")
        W("            if (v != 0.0) atomicAdd(&d.pairbuf[(size_t)pq * HFX_PAIRBUF_STRIDE + pe], v);")
        W("          }")
        W("          wave_sync();")
        W("        }")
        W("      }")
        W("    }")
        W("    if constexpr (!JONLY) {")
        # K in one pass (a outermost) or in two (K1 per a, K2 per b), per
        # class and precision as measured on C60 (KPASS_TWO)
        two64 = (cls, "double") in KPASS_TWO
        two32 = (cls, "float") in KPASS_TWO
        W("    if constexpr (sizeof(T) == sizeof(double) ? %s : %s) {" % (
            "false" if two64 else "true", "false" if two32 else "true"))
        # pass K: a outermost; X_ac, X_ad complete after each a, X_bc, X_bd
        # (nB rows, kept in registers) after the last a.  Rows of P for b are
        # reloaded per (a, b) from cache rather than held in registers.
        W("    {")
        W("      const T* bp = blk; HFX_PIN_GPTR(bp);")
        for ib in range(nB):
            zero("Kbc_%d" % ib, nC)
            zero("Kbd_%d" % ib, nD)
        for ia in range(nA):
            W("      {")
            zero("Kac", nC)
            zero("Kad", nD)
            for id_ in range(nD):
                W("      const T Pad_%d = PQ[(size_t)(aoA + %d) * nbf + aoD + %d];" % (id_, ia, id_))
            for ic in range(nC):
                W("      const T Pac_%d = PQ[(size_t)(aoA + %d) * nbf + aoC + %d];" % (ic, ia, ic))
            for ib in range(nB):
                W("      {")
                for id_ in range(nD):
                    W("      const T Pbd_%d = PQ[(size_t)(aoB + %d) * nbf + aoD + %d];" % (id_, ib, id_))
                for ic in range(nC):
                    W("      const T Pbc_%d = PQ[(size_t)(aoB + %d) * nbf + aoC + %d];" % (ic, ib, ic))
                for ic in range(nC):
                    for id_ in range(nD):
                        W("      { const T v = bp[%d]; Kac_%d = fma(v, Pbd_%d, Kac_%d); Kad_%d = fma(v, Pbc_%d, Kad_%d);"
                          " Kbc_%d_%d = fma(v, Pad_%d, Kbc_%d_%d); Kbd_%d_%d = fma(v, Pac_%d, Kbd_%d_%d); }" % (
                              lin(ia, ib, ic, id_), ic, id_, ic, id_, ic, id_,
                              ib, ic, id_, ib, ic, ib, id_, ic, ib, id_))
                pin("Kac", nC); pin("Kad", nD); pin("Kbc_%d" % ib, nC); pin("Kbd_%d" % ib, nD)
                W("      }")
            flush_row("Kac", nC, "aoA", ia, ia, "aoC", "wk")
            flush_row("Kad", nD, "aoA", ia, ia, "aoD", "wk")
            W("      }")
        for ib in range(nB):
            flush_row("Kbc_%d" % ib, nC, "aoB", ib, nA + ib, "aoC", "wk")
            flush_row("Kbd_%d" % ib, nD, "aoB", ib, nA + ib, "aoD", "wk")
        W("    }")
        W("    } else {")
        # pass K1
        W("    {")
        W("      const T* bp = blk; HFX_PIN_GPTR(bp);")
        pload("Pbd", "aoB", nB, "aoD", nD)
        pload("Pbc", "aoB", nB, "aoC", nC)
        for ia in range(nA):
            W("      {")
            zero("Kac", nC)
            zero("Kad", nD)
            for ib in range(nB):
                for ic in range(nC):
                    for id_ in range(nD):
                        W("      { const T v = bp[%d]; Kac_%d = fma(v, Pbd_%d_%d, Kac_%d); Kad_%d = fma(v, Pbc_%d_%d, Kad_%d); }" % (
                            lin(ia, ib, ic, id_), ic, ib, id_, ic, id_, ib, ic, id_))
                if (ib + 1) % gs == 0 or ib == nB - 1:
                    pin("Kac", nC); pin("Kad", nD)
            flush_row("Kac", nC, "aoA", ia, ia, "aoC", "wk")
            flush_row("Kad", nD, "aoA", ia, ia, "aoD", "wk")
            W("      }")
        W("    }")
        # pass K2
        W("    {")
        W("      const T* bp = blk; HFX_PIN_GPTR(bp);")
        pload("Pad", "aoA", nA, "aoD", nD)
        pload("Pac", "aoA", nA, "aoC", nC)
        for ib in range(nB):
            W("      {")
            zero("Kbc", nC)
            zero("Kbd", nD)
            for ia in range(nA):
                for ic in range(nC):
                    for id_ in range(nD):
                        W("      { const T v = bp[%d]; Kbc_%d = fma(v, Pad_%d_%d, Kbc_%d); Kbd_%d = fma(v, Pac_%d_%d, Kbd_%d); }" % (
                            lin(ia, ib, ic, id_), ic, ia, id_, ic, id_, ia, ic, id_))
                if (ia + 1) % gs == 0 or ia == nA - 1:
                    pin("Kbc", nC); pin("Kbd", nD)
            flush_row("Kbc", nC, "aoB", ib, nA + ib, "aoC", "wk")
            flush_row("Kbd", nD, "aoB", ib, nA + ib, "aoD", "wk")
            W("      }")
        W("    }")
        W("    }")
        W("    }")
        W("  }")
        W("  if (rows && curb >= 0) flush();")
        W("}")
        for multi in ("false", "true"):
            W("template __global__ void dig_%s<false, double, %s>(const DigestArgs);" % (cls, multi))
            W("template __global__ void dig_%s<true, double, %s>(const DigestArgs);" % (cls, multi))
            W("template __global__ void dig_%s<false, float, %s>(const DigestArgs);" % (cls, multi))

    def kernel_digest1(self, W):
        """One-pass digestion of a block class (single molecule, off-diagonal
        J and K builds): every element of a block is loaded once and used in
        all six contractions.  A workgroup of DIG1_NT threads walks a
        contiguous range of the bra-contiguous list in steps that never cross
        a bra pair (a step is the run of the current bra's quartets, at most
        DIG1_NT), so every quartet of a step has the resident bra; the rows of
        X (accumulated) and of P (read) of shells A and B sit in LDS.  Only
        P_cd is gathered from global memory.  Per quartet, with the outer
        shell O of the bra (the larger of A, B) looped at run time and the
        inner one I unrolled:
          J:  X_ab += 2w sum_cd e P_cd per (a, b) (summed over the wave into
              LDS), X_cd += 2w sum_ab e P_ab (staged to the pair buffer);
          K:  per o: X_oc -= w/2 sum_{i,d} e P_id, X_od -= w/2 sum_{i,c} e P_ic;
              held over o: X_ic -= w/2 sum_{o,d} e P_od, X_id -= w/2 sum_{o,c} e P_oc."""
        cls = self.cls
        nA, nB, nC, nD = self.dims()
        ne = nA * nB * nC * nD
        ncd = nC * nD
        oa = "b" if self.swap_bra else "a"
        ob = "a" if self.swap_bra else "b"
        oc = "b" if self.swap_ket else "a"
        od = "a" if self.swap_ket else "b"
        a_outer = nA >= nB
        n_o, n_i = (nA, nB) if a_outer else (nB, nA)
        # element stride of the outer index, and the LDS row of outer index o / inner index i
        o_stride = (nB * ncd if a_outer else ncd) * 64

        def lin_i(i, c, d):       # element offset of (o = 0, i, c, d)
            ia, ib = (0, i) if a_outer else (i, 0)
            return (((ia * nB + ib) * nC + c) * nD + d) * 64

        orow = "o" if a_outer else "%d + o" % nA
        irow = (lambda i: "%d" % (nA + i)) if a_outer else (lambda i: "%d" % i)

        nt32 = DIG1_NT32.get(cls, 512)
        assert nt32 % 64 == 0 and 64 <= nt32 <= 1024
        W("template <typename T>")
        W("__global__ void __launch_bounds__(sizeof(T) == sizeof(double) ? DIG1_NT : %d) dig1_%s(const DigestArgs d) {"
          % (nt32, cls))
        W("  extern __shared__ char hfx_smem[];")
        W("  constexpr int NR = %d;   // rows of A then rows of B" % (nA + nB))
        W("  double* srow = reinterpret_cast<double*>(hfx_smem);                 // [NR][nb] of X")
        W("  T* prow = reinterpret_cast<T*>(hfx_smem + d.prow_offset);           // [NR][nb] of P (P32 for FP32 blocks)")
        W("  int* wrun = reinterpret_cast<int*>(hfx_smem + d.run_offset);       // per wave: first lane past the bra's run")
        W("  const int qend = min(d.q1, *d.count);")
        W("  if (qend <= d.q0) return;")
        W("  const int nbf = d.nb;")
        W("  const int lane = threadIdx.x & (HFX_WAVE - 1);")
        W("  const int nwave = blockDim.x / HFX_WAVE;")
        W("  double* X = d.X;")
        W("  const T* PA;")
        W("  if constexpr (sizeof(T) == sizeof(float)) PA = reinterpret_cast<const T*>(d.P32);")
        W("  else PA = reinterpret_cast<const T*>(d.P);")
        W("  const double thq = d.thr_mol ? d.thr_mol[0] : d.thr;")
        W("  // this workgroup's contiguous range of the chunk")
        W("  const long long ntot = qend - d.q0;")
        W("  const int Q0 = d.q0 + (int)(ntot * blockIdx.x / gridDim.x);")
        W("  const int Q1 = d.q0 + (int)(ntot * (blockIdx.x + 1) / gridDim.x);")
        W("  for (int e = threadIdx.x; e < NR * nbf; e += blockDim.x) srow[e] = 0.0;")
        W("  __syncthreads();")
        W("  int curb = -1, rA = 0, rB = 0;")
        W("  for (int q = Q0; q < Q1;) {")
        W("    // the step: the run of list[q].x's quartets from q, at most blockDim.x")
        W("    const int bra = d.list[q].x;")
        W("    const int t = q + (int)threadIdx.x;")
        W("    const int2 bkg = d.list[min(t, Q1 - 1)];   // the lane's quartet, if the step reaches it")
        W("    const bool differ = t >= Q1 || bkg.x != bra;")
        # its ket side (pair record, then P_cd) is loaded here, so that the
        # dependent loads overlap the step's barriers and the bra's row
        # changes; a lane past the step keeps these values (it contributes 0)
        W("    const PairRec& pk = d.pairs[bkg.y];")
        W("    const int shC = pk.sh%s, shD = pk.sh%s;" % (oc, od))
        W("    const int aoC = pk.ao%s, aoD = pk.ao%s;" % (oc, od))
        for c in range(nC):
            W("    " + " ".join("const T Pcd_%d_%d = PA[(size_t)(aoC + %d) * nbf + aoD + %d];" % (c, dd, c, dd)
                                for dd in range(nD)))
        W("    const unsigned long long dm = __ballot(differ);")
        W("    if (lane == 0) wrun[threadIdx.x / HFX_WAVE] = dm ? (int)(threadIdx.x - lane) + __ffsll((long long)dm) - 1"
          " : (int)blockDim.x;")
        W("    __syncthreads();")
        W("    int n = (int)blockDim.x;")
        W("    for (int w = 0; w < nwave; ++w) n = min(n, wrun[w]);")
        W("    if (bra != curb) {")
        W("      if (curb >= 0) {   // X rows of the previous bra to X, row-coalesced")
        W("        for (int e = threadIdx.x; e < NR * nbf; e += blockDim.x) {")
        W("          const double v = srow[e];")
        W("          if (v != 0.0) {")
        W("            const int r = e / nbf, col = e - r * nbf;")
        W("            atomicAdd(&X[(size_t)(r < %d ? rA + r : rB + (r - %d)) * nbf + col], v);" % (nA, nA))
        W("            srow[e] = 0.0;")
        W("          }")
        W("        }")
        W("      }")
        W("      curb = bra;")
        W("      rA = d.pairs[bra].ao%s; rB = d.pairs[bra].ao%s;" % (oa, ob))
        W("      for (int e = threadIdx.x; e < NR * nbf; e += blockDim.x) {")
        W("        const int r = e / nbf, col = e - r * nbf;")
        W("        prow[e] = PA[(size_t)(r < %d ? rA + r : rB + (r - %d)) * nbf + col];" % (nA, nA))
        W("      }")
        W("    }")
        W("    __syncthreads();   // rows loaded / flushed, wrun read by every thread")
        W("    const int it = q + (int)(threadIdx.x - lane);   // first quartet of the wave")
        W("    q += n;")
        W("    if (it >= q) continue;   // the whole wave is past the step")
        W("    const bool valid = t < q;")
        W("    const int iq = valid ? t : it;")
        W("    const int slot = iq - d.q0;")
        W("    double vmax = d.vmax[slot];")
        W("    for (int c = 1; c < d.ncombo; ++c) vmax = fmax(vmax, d.vmax[(size_t)c * d.vstride + slot]);")
        W("    const bool act = valid && vmax >= thq;")
        W("    if (__ballot(act) == 0) continue;")
        W("    const T* blk = reinterpret_cast<const T*>(d.eblk) + (size_t)(slot >> 6) * %d + (slot & 63);" % (ne * 64))
        W("    const PairRec& pb = d.pairs[bra];   // every valid lane's bra")
        W("    const double w = act ? ((pb.sha == pb.shb) ? 0.5 : 1.0) * ((shC == shD) ? 0.5 : 1.0) : 0.0;")
        W("    const double wj = 2.0 * w, wk = -0.5 * w;")
        W("    " + " ".join("T Jcd_%d = 0;" % k for k in range(ncd)))

        def core(ki_slice, prefetch):
            """The contractions of one quartet.  ki_slice: X_ic, X_id go to the
            LDS rows after every (o, i) slice (fewer registers: the FP64 form
            holds no more than the running slice's sums); else they are summed
            over o in registers and flushed once (the FP32 form).  prefetch:
            the loads of slice i + 1 are issued before slice i is contracted
            (two slices of the block in flight per wave)."""

            def sload(i):
                W("          " + " ".join("s%d_%d = bp[%d];" % (i, c * nD + dd, lin_i(i, c, dd))
                                          for c in range(nC) for dd in range(nD)))
            if not ki_slice:
                for i in range(n_i):
                    W("      " + " ".join("T Kic_%d_%d = 0;" % (i, c) for c in range(nC)) + " "
                      + " ".join("T Kid_%d_%d = 0;" % (i, dd) for dd in range(nD)))
            W("      #pragma unroll 1")
            W("      for (int o = 0; o < %d; ++o) {" % n_o)
            W("        const T* bp = blk + (size_t)o * %d; HFX_PIN_GPTR(bp);" % o_stride)
            W("        const T* po = prow + (size_t)(%s) * nbf;   // P row of the outer index" % orow)
            for dd in range(nD):
                W("        const T Pod_%d = po[aoD + %d];" % (dd, dd))
            for c in range(nC):
                W("        const T Poc_%d = po[aoC + %d];" % (c, c))
            W("        " + " ".join("T Koc_%d = 0;" % c for c in range(nC)) + " "
              + " ".join("T Kod_%d = 0;" % dd for dd in range(nD)))
            if prefetch:
                W("        " + " ".join("T s0_%d = bp[%d];" % (c * nD + dd, lin_i(0, c, dd))
                                        for c in range(nC) for dd in range(nD)))
                for i in range(1, n_i):
                    W("        " + " ".join("T s%d_%d;" % (i, k) for k in range(ncd)))
            for i in range(n_i):
                W("        {")
                if prefetch and i + 1 < n_i:
                    sload(i + 1)
                W("          const T* pi = prow + (size_t)%s * nbf;   // P row of inner index %d" % (irow(i), i))
                for dd in range(nD):
                    W("          const T Pid_%d = pi[aoD + %d];" % (dd, dd))
                for c in range(nC):
                    W("          const T Pic_%d = pi[aoC + %d];" % (c, c))
                # P_ab of (a, b) = (o, i) or (i, o): row a of A, column b of B
                if a_outer:
                    W("          const T pab = prow[(size_t)o * nbf + rB + %d];" % i)
                else:
                    W("          const T pab = prow[(size_t)%d * nbf + rB + o];" % i)
                W("          T jab = 0;")
                if ki_slice:
                    W("          " + " ".join("T Kic_%d_%d = 0;" % (i, c) for c in range(nC)) + " "
                      + " ".join("T Kid_%d_%d = 0;" % (i, dd) for dd in range(nD)))
                for c in range(nC):
                    for dd in range(nD):
                        k = c * nD + dd
                        src = ("s%d_%d" % (i, k)) if prefetch else ("bp[%d]" % lin_i(i, c, dd))
                        W("          { const T v = %s; jab = fma(v, Pcd_%d_%d, jab); Jcd_%d = fma(v, pab, Jcd_%d);"
                          " Koc_%d = fma(v, Pid_%d, Koc_%d); Kod_%d = fma(v, Pic_%d, Kod_%d);"
                          " Kic_%d_%d = fma(v, Pod_%d, Kic_%d_%d); Kid_%d_%d = fma(v, Poc_%d, Kid_%d_%d); }" % (
                              src, c, dd, k, k, c, dd, c, dd, c, dd,
                              i, c, dd, i, c, i, dd, c, i, dd))
                # X_ab of this (a, b): summed over the wave (every lane has the resident bra)
                W("          double r = wj * (double)jab;")
                W("          r = wave_sum(r);")
                if a_outer:
                    W("          if (lane == 0 && r != 0.0) atomicAdd(&srow[(size_t)o * nbf + rB + %d], r);" % i)
                else:
                    W("          if (lane == 0 && r != 0.0) atomicAdd(&srow[(size_t)%d * nbf + rB + o], r);" % i)
                held = [] if ki_slice else (["Kic_%d_%d" % (i, c) for c in range(nC)]
                                            + ["Kid_%d_%d" % (i, dd) for dd in range(nD)])
                W("          " + " ".join("HFX_PIN(%s);" % n for n in ["Jcd_%d" % k for k in range(ncd)]
                                          + ["Koc_%d" % c for c in range(nC)] + ["Kod_%d" % dd for dd in range(nD)]
                                          + held) + " HFX_LOAD_FENCE();")
                if ki_slice:
                    kflush_i(i, "          ")
                W("        }")
            W("        if (act) {")
            W("          double* so = srow + (size_t)(%s) * nbf;" % orow)
            for c in range(nC):
                W("          atomicAdd(&so[aoC + %d], wk * (double)Koc_%d);" % (c, c))
            for dd in range(nD):
                W("          atomicAdd(&so[aoD + %d], wk * (double)Kod_%d);" % (dd, dd))
            W("        }")
            W("      }")
            if not ki_slice:
                for i in range(n_i):
                    kflush_i(i, "      ")

        def kflush_i(i, ind):
            W(ind + "if (act) {")
            W(ind + "  double* si = srow + (size_t)%s * nbf;" % irow(i))
            for c in range(nC):
                W(ind + "  atomicAdd(&si[aoC + %d], wk * (double)Kic_%d_%d);" % (c, i, c))
            for dd in range(nD):
                W(ind + "  atomicAdd(&si[aoD + %d], wk * (double)Kid_%d_%d);" % (dd, i, dd))
            W(ind + "}")

        W("    if constexpr (sizeof(T) == sizeof(double)) {")
        core(True, cls in DIG1_PREFETCH64)
        W("    } else {")
        core(cls in DIG1_KSLICE32, DIG1_PREFETCH32)
        W("    }")
        # X_cd through the wave's LDS staging into the ket pair's slot of the pair buffer
        W("    {")
        W("      constexpr int TG = HFX_WAVE < 16 ? HFX_WAVE : 16;")
        W("      double* tb = reinterpret_cast<double*>(hfx_smem + d.tb_offset) + (threadIdx.x / HFX_WAVE) * (TG * %d);" % ncd)
        W("      const double sc = act ? wj : 0.0;")
        W("      for (int r = 0; r < HFX_WAVE / TG; ++r) {")
        W("        if (lane / TG == r) {")
        W("          double* tt = tb + (lane - r * TG) * %d;" % ncd)
        W("          " + " ".join("tt[%d] = sc * (double)Jcd_%d;" % (k, k) for k in range(ncd)))
        W("        }")
        W("        wave_sync();")
        W("        for (int f0 = 0; f0 < TG * %d; f0 += HFX_WAVE) {" % ncd)
        W("          const int f = f0 + lane;")
        W("          const bool ok = f < TG * %d;" % ncd)
        W("          const int ql = ok ? f / %d : 0, e = ok ? f - ql * %d : 0;" % (ncd, ncd))
        W("          const int pq = __shfl(bkg.y, r * TG + ql);")
        W("          const int pe = %s;" % ("(e %% %d) * %d + e / %d" % (nD, nC, nD) if self.swap_ket else "e"))
        W("          const double v = ok ? tb[f] : 0.0;")
        W("          if (v != 0.0) atomicAdd(&d.pairbuf[(size_t)pq * HFX_PAIRBUF_STRIDE + pe], v);")
        W("        }")
        W("        wave_sync();")
        W("      }")
        W("    }")
        W("  }")
        W("  __syncthreads();")
        W("  if (curb >= 0) {")
        W("    for (int e = threadIdx.x; e < NR * nbf; e += blockDim.x) {")
        W("      const double v = srow[e];")
        W("      if (v != 0.0) {")
        W("        const int r = e / nbf, col = e - r * nbf;")
        W("        atomicAdd(&X[(size_t)(r < %d ? rA + r : rB + (r - %d)) * nbf + col], v);" % (nA, nA))
        W("      }")
        W("    }")
        W("  }")
        W("}")
        W("template __global__ void dig1_%s<double>(const DigestArgs);" % cls)
        W("template __global__ void dig1_%s<float>(const DigestArgs);" % cls)

    def kernel(self, split):
        L = []
        W = L.append
        ops_prim, ops_bra, ops_ctr = Ops(), Ops(), Ops()
        W("// (%s|%s) generated by hfcode_amd/gen/gen_eri.py -- do not edit" % (self.cls[:2], self.cls[2:]))
        W("// build centres: A=%s B=%s C=%s D=%s; Boys m <= %d; %s" % (
            self.A, self.B, self.C, self.D, self.M, "block + k_digest" if split else "fused digestion"))
        W("#include \"hfx_device.h\"")
        if split:
            W("#include \"hfx_kernels.h\"")
        W("namespace hfx {")
        ex_prim, ex_bra, ex_ctr = Ops(), Ops(), Ops()
        if split:
            # algorithmic operation counts: one Boys evaluation and one vertical
            # recursion per primitive quartet over the union of the class's
            # targets (every part combination in one pass), whatever the
            # grouping of the emitted kernels, which repeats the Boys
            # evaluation and the shared recursion steps once per kernel
            union = [[(bp, kp) for bp in self.bra_parts for kp in self.ket_parts]]
            self.kernel_block(lambda _line: None, ops_prim, ops_bra, ops_ctr, plan=union)
            cb64 = self.kernel_block(W, ex_prim, ex_bra, ex_ctr)
            cb32 = self.kernel_block(W, Ops(), Ops(), Ops(), fp32=True)
            nA, nB, nC, nD = self.dims()
            # digestion in dig_<class>: 6 contractions of every element + block max
            ops_ctr.n += 13 * nA * nB * nC * nD
            ex_ctr.n += 13 * nA * nB * nC * nD
        else:
            self.kernel_fused(W, ops_prim, ops_bra, ops_ctr)
            ex_prim.n, ex_bra.n, ex_ctr.n = ops_prim.n, ops_bra.n, ops_ctr.n
            W("template __global__ void eri_%s<double, false>(const ClassArgs);" % self.cls)
            W("template __global__ void eri_%s<double, true>(const ClassArgs);" % self.cls)
            W("template __global__ void eri_%s<float, false>(const ClassArgs);" % self.cls)
        W("}  // namespace hfx")
        info = {
            "class": self.cls, "M": self.M, "bra": self.bra_type, "ket": self.ket_type,
            "split": bool(split),
            "combos": len(self.group_plan(True)) if split else 1,
            "combos32": len(self.group_plan(True, True)) if split else 0,
            "parts": len(self.bra_parts) * len(self.ket_parts),
            "flops_prim_vrr": ops_prim.n,
            "flops_braprim": ops_bra.n,
            "flops_quartet": ops_ctr.n,
            "exec_prim_vrr": ex_prim.n,
            "exec_braprim": ex_bra.n,
            "exec_quartet": ex_ctr.n,
            "all_prims": self.all_prims,
        }
        if split:
            # pre-kernels per combination kernel (split combinations) and the
            # contracted values they store per quartet
            info["npre"] = [len(self.split_parts(g) or [0]) - 1 for g in self.group_plan(True)]
            info["npre32"] = [len(self.split_parts(g, True) or [0]) - 1 for g in self.group_plan(True, True)]
            info["cbuf64"], info["cbuf32"] = cb64, cb32
        return "\n".join(L) + "\n", info

    def digest_source(self):
        """Translation unit of the class's digestion kernel (block classes)."""
        L = []
        W = L.append
        W("// digestion of (%s|%s) blocks, generated by hfcode_amd/gen/gen_eri.py -- do not edit" % (
            self.cls[:2], self.cls[2:]))
        W("#include \"hfx_device.h\"")
        W("#include \"hfx_kernels.h\"")
        W("namespace hfx {")
        self.kernel_digest(W, None)
        self.kernel_digest1(W)
        W("}  // namespace hfx")
        return "\n".join(L) + "\n"

    def group_costs(self, fp32=False):
        """Relative cost of each emitted combination kernel of a block class:
        its operations per primitive quartet + per bra primitive pair / 3 + per
        quartet / 8 (about 8 primitive quartets per quartet and 3 bra
        primitive pairs in split-valence bases).  Used to spread a chunk's
        kernels over streams."""
        snaps = []
        self.kernel_block(lambda _line: None, Ops(), Ops(), Ops(), fp32=fp32, group_ops=snaps)
        out, prev = [], (0, 0, 0)
        for sn in snaps:
            d = [a - b for a, b in zip(sn, prev)]
            out.append(d[0] + d[1] / 3.0 + d[2] / 8.0)
            prev = sn
        return out

    def fused_cost(self):
        """Doubles a fused kernel keeps live: accumulators + J/K partial sums + density blocks."""
        def nc(ls):
            return sum((l + 1) * (l + 2) // 2 for l in ls)
        e_all = sorted({l for bp in self.bra_parts for l in self.erange(bp)[2]})
        acc = sum(nc(self.erange(bp)[2]) * nc(self.frange(kp)[2]) for bp in self.bra_parts for kp in self.ket_parts)
        tmp = sum(nc(e_all) * nc(self.frange(kp)[2]) for kp in self.ket_parts) if self.has_L_bra else 0
        nA, nB, nC, nD = self.dims()
        dig = 2 * (nA * nB + nC * nD + nA * nC + nA * nD + nB * nC + nB * nD)
        return acc + tmp + dig


SPLIT_COST = 200    # classes whose fused kernel would keep more doubles live are split
# Block classes (FP64 kernels, FP32 kernels) whose largest part combinations
# are split over several kernels by ket component (Gen.split_parts) when
# their contracted [e0|f0] exceed SPLIT_ACC doubles (2 x SPLIT_ACC floats):
# pre-kernels store their parts, the last kernel runs the HRR one ket output
# at a time.  Measured on the C60 build (class spans, ms, unsplit / split at
# 128): (dd|ll) FP64 10.9 / 9.8, but (dl|dl) 22.2 / 24.1, (dd|dl) 18.8 / 20.8,
# (dd|dd) 2.7 / 5.2, and the FP32 kernels of (dd|ll) 2.3 / 3.7 -- the last
# kernel's loads of the stored parts and the repeated recursions cost more
# than the spills they remove, once spilling kernels keep their scratch
# memory (raise_scratch_limit in hfx_host.hip).
SPLIT_ACC = 128
SPLIT_CLASSES = {"ddll": (True, False)}

# Digestion classes and precisions whose K contractions keep two passes over
# the block (K1 per a, K2 per b) instead of one: measured on the C60 /
# 6-31G(d) build, the one-pass form took (dl|dl) FP64 from 32.6 to 38.8 ms
# (260 registers: one wave per SIMD), (dd|ll) FP64 from 16.0 to 16.8 ms and
# (dl|ll) FP32 from 16.3 to 17.1 ms (15.7 to 16.5 ms again with the 10.5 GB
# integral buffer of a lone context); it is faster for the other classes
# (e.g. (dl|ll) FP64 39.9 -> 36.9 ms, (dd|ll) FP32 4.2 -> 2.8 ms).
KPASS_TWO = {("dldl", "double"), ("ddll", "double"), ("dlll", "float")}
# Block classes whose one-pass digestion (dig1_<class>) measured faster than
# the multi-pass one on the C60 / 6-31G(d) builds, per precision of the list
# (bit 0 FP64, bit 1 FP32); the library also requires long bra runs.  Class
# spans, ms, one-pass / multi-pass (FP64 form flushing X_ic, X_id per slice):
# FP64 build (dl|ll) 26.9 / 34.5, (dl|dl) 26.4 / 31.8, (dd|dl) 21.5 / 22.8,
# (ll|ll) 19.4 / 21.8, (dd|ll) 12.2 / 14.7, but (sl|ll) 8.8 / 8.5, (sl|dl)
# 7.3 / 6.7, (dd|dd) 3.4 / 3.2; FP32 lists of the mixed build (dl|ll) 12.3 /
# 16.7, (ll|ll) 9.5 / 9.8, (dd|dd) 1.13 / 1.27, but (dd|dl) 6.8 / 4.8, (sl|ll)
# 5.3 / 4.9, (sl|dl) 4.7 / 3.7.
# The FP64 list of the mixed build has shorter runs (mean quartets per bra
# pair: (dl|ll) 1431, (dd|dl) 831, (ll|ll) 803, (dl|dl) 385): there (dd|dl)
# still gains (4.4 / 5.2 ms) while (ll|ll) loses (7.7 / 7.1), so the minimum
# mean run is per class (DIG1_MIN_RUN, default 1024 quartets).
# Re-measured with the wave sums of X_ab on DPP (no LDS shuffles) and global
# block loads, one-pass / table: (sl|dl) FP64 build 5.9 / 6.5, mixed FP32
# 3.02 / 3.38; (sl|dd) FP64 1.59 / 1.81, mixed FP64 / FP32 0.42 / 0.53,
# 0.85 / 1.14 (mean runs 244 / 308); (dd|dd) FP32 0.91 / 1.10 (mean run
# 193); (sd|dl) mixed FP64 0.35 / 0.40 (mean run 841); (sl|ll) still loses
# (8.55 / 8.33, 4.73 / 4.58).
DIG1_USE = {"dlll": 3, "llll": 3, "ddll": 3, "sdll": 3, "sddl": 3, "dldl": 1, "dddl": 1, "dddd": 2, "sddd": 2,
            "sldl": 3, "sldd": 3}
DIG1_MIN_RUN = {"dddl": 512, "sldl": 512, "sddl": 512, "sldd": 128, "dddd": 128, "sddd": 128}
# one-pass digestion: next slice's loads issued before the current slice's
# contractions, for the FP32 form and the FP64 form of the classes listed
# (measured, C60 builds, ms: FP32 (dl|ll) 12.3 -> 11.9, (ll|ll) 9.3 -> 9.1,
# (dd|ll) 2.39 -> 2.27; FP64 (ll|ll) 19.2 -> 18.8; the FP64 forms of (dl|dl)
# and (dd|dl) would spill 150 bytes per lane)
DIG1_PREFETCH32 = True
DIG1_PREFETCH64 = {"dlll", "llll"}
# FP32 one-pass digestion: workgroup size per class (default 512 threads, as
# the FP64 form), and the classes whose FP32 form flushes X_ic, X_id per slice
DIG1_NT32 = {}
DIG1_KSLICE32 = set()
FP64_GROUP_ACC = 64    # block kernels: contracted accumulators per kernel, at most (FP64)
FP32_GROUP_ACC = 128   # the same for the FP32 block kernels
# per-class budgets picked by measurement (hfcode_amd/gen/autotune_groups.py);
# classes not in the table use the defaults above
GROUP_BUDGETS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "group_budgets.json")


def load_group_budgets(path=GROUP_BUDGETS_FILE):
    """{class: [fp64 budget, fp32 budget]} from the tuned table, {} without one."""
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        return {k: tuple(v) for k, v in json.load(fh)["budgets"].items()}


GROUP_BUDGETS = load_group_budgets()
DIG_SLICE_PAIR_MAX = 16   # digestion: two block slices (nC x nD elements) per compiler fence when nC nD <= this


def boys_flops(M, K):
    """Table-Taylor Boys: K-1 products for the Taylor factors, then K-1 fma + 1
    mul per order; plus index/offset arithmetic (3)."""
    return (K - 1) + (M + 1) * (2 * (K - 1) + 1) + 3


# Algorithmic operation counts per class, frozen: the FLOPs a build is priced
# at (bench.py's roofline) come from this committed table, not from the code
# the generator currently emits, so kernel tuning cannot move them.
# Written by ``python -m hfcode_amd.gen.gen_eri --freeze``; a CPU test checks
# that the generator's union model still reproduces it.
FLOPS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "flops_per_class.json")
FROZEN_KEYS = ("flops_prim", "flops_prim_fp32", "flops_braprim", "flops_quartet")


def algorithmic_counts():
    """{class: {flops_prim, flops_prim_fp32, flops_braprim, flops_quartet, M,
    exec_prim, exec_braprim, exec_quartet}} from the generator's model."""
    out = {}
    for cls in CLASSES:
        g = Gen(cls)
        split = g.fused_cost() > SPLIT_COST
        _, info = g.kernel(split)
        ex_boys = sum(boys_flops(m, BOYS_TERMS["double"]) for m in (
            [g.group_M(grp) for grp in g.group_plan(True)] if split else [info["M"]]))
        out[cls] = {
            "M": info["M"], "split": bool(split),
            "flops_prim": info["flops_prim_vrr"] + boys_flops(info["M"], BOYS_TERMS["double"]),
            "flops_prim_fp32": info["flops_prim_vrr"] + boys_flops(info["M"], BOYS_TERMS["float"]),
            "flops_braprim": info["flops_braprim"], "flops_quartet": info["flops_quartet"],
            "exec_prim": info["exec_prim_vrr"] + ex_boys, "exec_braprim": info["exec_braprim"],
            "exec_quartet": info["exec_quartet"],
        }
    return out


def load_frozen(path=FLOPS_FILE):
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)["classes"]


def freeze(path=FLOPS_FILE):
    counts = algorithmic_counts()
    doc = {
        "about": "Algorithmic operations per class (add = mul = 1, fma = 2): flops_prim per primitive "
                 "quartet (one Boys evaluation and one vertical recursion over the union of the class's "
                 "targets, contraction into the accumulators), flops_braprim per bra primitive pair, "
                 "flops_quartet per contracted quartet (horizontal recursion, spherical transform, six J/K "
                 "contractions and the block maximum of every element).  exec_* count what the emitted "
                 "kernels execute (each kernel of a block class repeats the Boys evaluation and the shared "
                 "recursion steps); they are informational.",
        "classes": counts,
    }
    with open(path, "w") as fh:
        json.dump(doc, fh, indent=1, sort_keys=True)
        fh.write("\n")


def main(outdir):
    os.makedirs(outdir, exist_ok=True)
    frozen = load_frozen()
    table = {}
    names = []
    for cls in CLASSES:
        g = Gen(cls)
        split = g.fused_cost() > SPLIT_COST
        src, info = g.kernel(split)
        path = os.path.join(outdir, "eri_%s.hip" % cls)
        dpath = os.path.join(outdir, "dig_%s.hip" % cls)
        if split:
            dsrc = g.digest_source()
            if not os.path.exists(dpath) or open(dpath).read() != dsrc:
                with open(dpath, "w") as fh:
                    fh.write(dsrc)
        elif os.path.exists(dpath):
            os.remove(dpath)
        old = open(path).read() if os.path.exists(path) else None
        if old != src:
            with open(path, "w") as fh:
                fh.write(src)
        # algorithmic: one Boys evaluation per primitive quartet; executed: one per
        # emitted FP64 kernel (group of part combinations) of the class
        info["flops_prim"] = info["flops_prim_vrr"] + boys_flops(info["M"], BOYS_TERMS["double"])
        info["flops_prim_fp32"] = info["flops_prim_vrr"] + boys_flops(info["M"], BOYS_TERMS["float"])
        info["exec_prim"] = info["exec_prim_vrr"] + sum(
            boys_flops(m, BOYS_TERMS["double"]) for m in ([g.group_M(grp) for grp in g.group_plan(True)]
                                                          if split else [info["M"]]))
        if frozen is not None:
            for k in FROZEN_KEYS:
                info[k] = frozen[cls][k]
        info["dims"] = g.dims()
        info["combo_M"] = [g.group_M(grp) for grp in g.group_plan(True)] if split else []
        info["combo_M32"] = [g.group_M(grp) for grp in g.group_plan(True, True)] if split else []
        info["combo_cost"] = g.group_costs() if split else []
        info["combo_cost32"] = g.group_costs(True) if split else []
        info["swap_bra"] = g.swap_bra
        info["swap_ket"] = g.swap_ket
        table[cls] = info
        names.append(cls)
    D = ["// generated by hfcode_amd/gen/gen_eri.py -- do not edit",
         "#pragma once", "#include \"hfx_device.h\"", "#include \"hfx_kernels.h\"", "namespace hfx {"]
    for c in names:
        if table[c]["split"]:
            for k in range(table[c]["combos"]):
                D.append("template <typename R> __global__ void eri_%s_c%d(const ClassArgs);" % (c, k))
            for k in range(table[c]["combos32"]):
                D.append("template <typename R> __global__ void eri_%s_f%d(const ClassArgs);" % (c, k))
            for kname, key in (("c", "npre"), ("f", "npre32")):
                for k, m in enumerate(table[c][key]):
                    for j in range(m):
                        D.append("template <typename R> __global__ void eri_%s_%s%dv%d(const ClassArgs);" % (c, kname, k, j))
            D.append("template <bool JONLY, typename T, bool MULTI> __global__ void dig_%s(const DigestArgs);" % c)
            D.append("template <typename T> __global__ void dig1_%s(const DigestArgs);" % c)
        else:
            D.append("template <typename R, bool JONLY> __global__ void eri_%s(const ClassArgs);" % c)
    D.append("struct ClassKernels {")
    D.append("  const void* f64; const void* f64j; const void* f32; int M; int all_prims;")
    D.append("  double flops_prim, flops_prim_fp32, flops_braprim, flops_quartet;")
    D.append("  int split; int dims[4]; int swap_bra, swap_ket;")
    D.append("  int ncombo; int combo_M[16]; const void* c64[16];")
    D.append("  int ncombo32; int combo_M32[16]; const void* c32[16];   // FP32: grouped combinations")
    D.append("  const void* dig; const void* digj; const void* digf;")
    D.append("  const void* digm; const void* digjm; const void* digfm;   // batch contexts (molecule-aware)")
    D.append("  float combo_cost[16]; float combo_cost32[16];   // relative cost of each combination kernel")
    D.append("  const void* dig1; const void* dig1f;   // one-pass digestion (one molecule, J and K), FP64 / FP32 blocks")
    D.append("  int dig1_use;   // bit 0 / 1: the one-pass digestion measured faster for FP64 / FP32 blocks")
    D.append("  int dig1_min_run;   // ... when a list's mean run of quartets per bra pair is at least this")
    D.append("  int npre[16]; const void* pre64[16][16];      // split combinations: kernels that run before c64[k]")
    D.append("  int npre32[16]; const void* pre32[16][16];    // ... before c32[k]")
    D.append("  int cbuf64, cbuf32;   // contracted values they store per quartet (doubles / floats)")
    D.append("  int dig1_nt32;   // threads per workgroup of the FP32 one-pass digestion (FP64: DIG1_NT)")
    D.append("};")
    D.append("static const ClassKernels kClassKernels[%d] = {" % len(names))
    for c in names:
        t = table[c]
        if t["split"]:
            fns = "nullptr, nullptr, nullptr"
            nc = t["combos"]
            cm = "{" + ", ".join(str(x) for x in t["combo_M"]) + "}"
            c64 = "{" + ", ".join("(const void*)&eri_%s_c%d<double>" % (c, k) for k in range(nc)) + "}"
            nc32 = t["combos32"]
            cm32 = "{" + ", ".join(str(x) for x in t["combo_M32"]) + "}"
            c32 = "{" + ", ".join("(const void*)&eri_%s_f%d<float>" % (c, k) for k in range(nc32)) + "}"
            dg = ("(const void*)&dig_%s<false, double, false>, (const void*)&dig_%s<true, double, false>, "
                  "(const void*)&dig_%s<false, float, false>, " % (c, c, c))
            dg += ("(const void*)&dig_%s<false, double, true>, (const void*)&dig_%s<true, double, true>, "
                   "(const void*)&dig_%s<false, float, true>" % (c, c, c))
            dg += ", {%s}, {%s}" % (", ".join("%.1ff" % x for x in t["combo_cost"]),
                                    ", ".join("%.1ff" % x for x in t["combo_cost32"]))
            dg += ", (const void*)&dig1_%s<double>, (const void*)&dig1_%s<float>, %d, %d" % (
                c, c, DIG1_USE.get(c, 0), DIG1_MIN_RUN.get(c, 1024))

            def pre(npre, kname, tname_):
                assert max(npre) <= 16, (c, npre)
                cnt = "{" + ", ".join(str(x) for x in npre) + "}"
                ptr = "{" + ", ".join("{" + ", ".join("(const void*)&eri_%s_%s%dv%d<%s>" % (c, kname, k, j, tname_)
                                                      for j in range(m)) + "}" if m else "{nullptr}"
                                      for k, m in enumerate(npre)) + "}"
                return cnt + ", " + ptr
            dg += ", %s, %s, %d, %d, %d" % (pre(t["npre"], "c", "double"), pre(t["npre32"], "f", "float"),
                                            t["cbuf64"], t["cbuf32"], DIG1_NT32.get(c, 512))
        else:
            fns = "(const void*)&eri_%s<double, false>, (const void*)&eri_%s<double, true>, (const void*)&eri_%s<float, false>" % (c, c, c)
            nc, cm, c64 = 0, "{0}", "{nullptr}"
            nc32, cm32, c32 = 0, "{0}", "{nullptr}"
            dg = ("nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, {0.0f}, {0.0f}, nullptr, nullptr, 0, 0, "
                  "{0}, {{nullptr}}, {0}, {{nullptr}}, 0, 0, 0")
        D.append("  {%s, %d, %d, %d, %d, %d, %d, %d, {%d, %d, %d, %d}, %d, %d, %d, %s, %s, %d, %s, %s, %s}," % (
            fns, t["M"], int(t["all_prims"]), t["flops_prim"], t["flops_prim_fp32"], t["flops_braprim"],
            t["flops_quartet"], int(t["split"]), *t["dims"], int(t["swap_bra"]), int(t["swap_ket"]),
            nc, cm, c64, nc32, cm32, c32, dg))
    D.append("};")
    D.append("}  // namespace hfx")
    path = os.path.join(outdir, "eri_dispatch.h")
    src = "\n".join(D) + "\n"
    old = open(path).read() if os.path.exists(path) else None
    if old != src:
        with open(path, "w") as fh:
            fh.write(src)
    with open(os.path.join(outdir, "flops_per_class.json"), "w") as fh:
        json.dump(table, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--freeze":
        freeze()
    else:
        main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(__file__), "..", "csrc", "gen"))
