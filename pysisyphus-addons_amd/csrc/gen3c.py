"""Generator of the register-resident 3c2e class kernels (int3c2e_gen_*.hip).

For a class (la, lb, lc) and a group of aux components m, it emits straight-line
device code that, per lane (= one auxiliary shell), per primitive pair of the
wave's shell pair and per auxiliary primitive:
  1. scales the Boys values g_N = s (-alpha)^N F_N (N <= L = la+lb+lc);
  2. evaluates the closed-form Hermite integrals R_tuv for lc <= t+u+v <= L,
        R_tuv = sum_i c_ti (2X)^{t-2i} sum_j c_uj (2Y)^{u-2j} sum_k c_vk (2Z)^{v-2k} g_{t+u+v-i-j-k}
     with the inner sums memoised (z[v][n], y[u][v][n]);
  3. scatters every R_tuv into W[h][m] = sum S^{lc}_m(tau) R_{h+tau} (Hobson: the
     aux function only touches Hermite order lc);
  4. contracts W with the wave-uniform bra coefficients
        acc[a b][m] += E^{ab}_{tuv} W[tuv][m]
     (E^{ab}_{tuv} = c_i c_j E^{ax bx}_t E^{ay by}_u E^{az bz}_v, read from the
     per-primitive-pair record, i.e. through the scalar cache).
After the primitive loops the Cartesian (a, b) indices are transformed to
normalized real solid harmonics and the block is stored.

Each class is emitted twice: int3c2e_gen_* in FP64 and int3c2e_gen32_* for the
mixed-precision path, where steps 2-4 run in FP32 (bra E-products from an FP32
copy of the records) while the Boys values, the prefactors and the sum over
primitive pairs stay FP64 (per primitive pair the FP32 accumulators are added
into FP64 ones).  Both take the aux-list window [nbeg, nsurv) of each pair.

Usage: python gen3c.py <outdir>   (writes int3c2e_gen_<la><lb>.hip and gen3c_classes.h)
"""

import math
import sys
from fractions import Fraction
from pathlib import Path

# classes generated: la <= LA_MAX, lb <= la, lc <= LC_MAX
LA_MAX = 2
LC_MAX = 3
ACC_LIMIT = 20     # Cartesian accumulators per lane (ncart_a ncart_b x m per group), FP64
ACC_LIMIT32 = 40   # the same for the FP32 variants (half the registers per value)
# FP64 classes compiled for 3 waves per SIMD (at most 168 VGPRs) instead of
# the 2 the compiler picks by itself; small spills at most.  Measured on the
# 60-atom s-d / aux s-f workload: 3c2e phase 4.36 -> 4.17 ms.  (dd|p) and the
# larger classes spill too much at 3 waves and stay as they are.
MIN_WAVES = {(1, 1, 3): 3, (2, 0, 2): 3, (2, 0, 3): 3, (2, 1, 2): 3, (2, 1, 3): 3, (2, 2, 0): 3}


def cart(l):
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def sph_coeffs(L):
    """Real solid harmonics (m = -L..L) in the Cartesian order of cart(L),
    scaled by 1/sqrt((2L-1)!!) (Helgaker, Jorgensen, Olsen, 6.4.47-6.4.50)."""
    out = []
    idx = {c: i for i, c in enumerate(cart(L))}
    for m in range(-L, L + 1):
        am = abs(m)
        vm2 = 1 if m < 0 else 0
        N = 1.0 / (2 ** am * math.factorial(L)) * math.sqrt(
            2.0 * math.factorial(L + am) * math.factorial(L - am) / (2.0 if m == 0 else 1.0))
        row = [Fraction(0)] * len(idx)
        for t in range((L - am) // 2 + 1):
            for u in range(t + 1):
                for v2 in range(vm2, am + 1, 2):
                    lx = 2 * t + am - 2 * u - v2
                    ly = 2 * u + v2
                    lz = L - 2 * t - am
                    sgn = -1 if ((t + (v2 - vm2) // 2) & 1) else 1
                    c = Fraction(sgn) * Fraction(1, 4 ** t) * math.comb(L, t) * math.comb(L - t, am + t) \
                        * math.comb(t, u) * math.comb(am, v2)
                    row[idx[(lx, ly, lz)]] += c
        scale = N / math.sqrt(dfact(2 * L - 1))
        out.append([float(c) * scale for c in row])
    return out


def herm_list(L):
    return [(t, u, s - t - u) for s in range(L + 1) for t in range(s, -1, -1) for u in range(s - t, -1, -1)]


def herm_index(t, u, v):
    s = t + u + v
    return s * (s + 1) * (s + 2) // 6 + (s - t) * (s - t + 1) // 2 + v


def ct(t, i):
    return math.factorial(t) // (math.factorial(i) * math.factorial(t - 2 * i))


def record_layout(la, lb):
    """Order of the E-products in a bra record (must match bra_records_kernel)."""
    items = []
    for ia, a in enumerate(cart(la)):
        for ib, b in enumerate(cart(lb)):
            for t in range(a[0] + b[0] + 1):
                for u in range(a[1] + b[1] + 1):
                    for v in range(a[2] + b[2] + 1):
                        items.append((ia, ib, t, u, v))
    return items


def fmt(x):
    r = repr(float(x))
    if "e" not in r and "." not in r and "inf" not in r and "nan" not in r:
        r += ".0"
    return r


def mgroups(la, lb, lc, limit=ACC_LIMIT):
    nab = len(cart(la)) * len(cart(lb))
    ncc = 2 * lc + 1
    mg = max(1, min(ncc, limit // nab))
    ng = -(-ncc // mg)
    mg = -(-ncc // ng)
    groups = []
    m = 0
    while m < ncc:
        groups.append(list(range(m, min(ncc, m + mg))))
        m += mg
    return groups


def fmtr(x, real):
    return fmt(x) + ("f" if real == "float" else "")


def gen_group(la, lb, lc, ms, real="double"):
    """Device function body for one group of aux components ms, in the
    arithmetic type real (double or float)."""
    L = la + lb + lc
    lab = la + lb
    H = herm_list(lab)
    hidx = {h: i for i, h in enumerate(H)}
    S = sph_coeffs(lc)
    f32 = real == "float"
    FMA = "fmaf" if f32 else "fma"
    lit = lambda x: fmtr(x, real)
    lines = []
    w = lines.append
    # W accumulators
    nW = len(H) * len(ms)
    w(f"      {real} W[{nW}];")
    w(f"#pragma unroll")
    w(f"      for (int q = 0; q < {nW}; q++) W[q] = {lit(0.0)};")
    # memoised inner sums z[v][n] = sum_k c_vk pz[v-2k] g[n-k]
    zdone, ydone = {}, {}

    def zname(v, n):
        if (v, n) not in zdone:
            terms = []
            for k in range(v // 2 + 1):
                c = ct(v, k)
                pz = f"pz{v - 2 * k}"
                terms.append(f"{lit(c)} * {pz} * g{n - k}" if c != 1 else f"{pz} * g{n - k}")
            name = f"z{v}_{n}"
            w(f"      const {real} {name} = {' + '.join(terms)};")
            zdone[(v, n)] = name
        return zdone[(v, n)]

    def yname(u, v, n):
        if (u, v, n) not in ydone:
            terms = []
            for j in range(u // 2 + 1):
                c = ct(u, j)
                zn = zname(v, n - j)
                terms.append(f"{lit(c)} * py{u - 2 * j} * {zn}" if c != 1 else f"py{u - 2 * j} * {zn}")
            name = f"y{u}_{v}_{n}"
            w(f"      const {real} {name} = {' + '.join(terms)};")
            ydone[(u, v, n)] = name
        return ydone[(u, v, n)]

    # S terms per m in group: list of (coef, tau)
    cl = cart(lc)
    sterms = []
    for m in ms:
        sterms.append([(S[m][i], cl[i]) for i in range(len(cl)) if S[m][i] != 0.0])
    # R_tuv for orders lc..L, consumed immediately; emitted grouped by v then u
    # so that the memoised partial sums z[v][.] and y[u][v][.] die early.
    order = []
    for v in range(L + 1):
        for u in range(L - v + 1):
            for t in range(L - u - v + 1):
                if t + u + v >= lc:
                    order.append((t, u, v))
    cur_v, cur_u = None, None
    for (t, u, v) in order:
        if v != cur_v:
            if cur_v is not None:
                w("      }")
                w("      }")
            w("      {")
            zdone.clear()
            cur_v, cur_u = v, None
            for n in range(v, L + 1):
                zname(v, n)
        if u != cur_u:
            if cur_u is not None:
                w("      }")
            w("      {")
            ydone.clear()
            cur_u = u
        consumers = []
        for mi, terms in enumerate(sterms):
            for (c, tau) in terms:
                h = (t - tau[0], u - tau[1], v - tau[2])
                if min(h) < 0 or sum(h) > lab:
                    continue
                consumers.append((hidx[h] * len(ms) + mi, c))
        if not consumers:
            continue
        terms = []
        for i in range(t // 2 + 1):
            c = ct(t, i)
            yn = yname(u, v, t + u + v - i)
            terms.append(f"{lit(c)} * px{t - 2 * i} * {yn}" if c != 1 else f"px{t - 2 * i} * {yn}")
        w(f"      {{ const {real} r = {' + '.join(terms)};")
        for q, c in consumers:
            w(f"        W[{q}] = {FMA}({lit(c)}, r, W[{q}]);")
        w("      }")
    if cur_v is not None:
        w("      }")
        w("      }")
    # contraction with the uniform bra record
    nb = len(cart(lb))
    acc, rec = ("accf", "rf") if f32 else ("acc", "rec")
    for k, (ia, ib, t, u, v) in enumerate(record_layout(la, lb)):
        h = hidx[(t, u, v)]
        w(f"      {{ const {real} e = {rec}[{4 + k}];")
        for mi in range(len(ms)):
            q = (ia * nb + ib) * len(ms) + mi
            w(f"        {acc}[{q}] = {FMA}(e, W[{h * len(ms) + mi}], {acc}[{q}]);")
        w("      }")
    return lines


def gen_store(la, lb, lc, ms):
    """Cartesian -> spherical of (a, b) and store of the group's columns."""
    Sa, Sb = sph_coeffs(la), sph_coeffs(lb)
    na, nb = len(cart(la)), len(cart(lb))
    lines = []
    w = lines.append
    for mu in range(2 * la + 1):
        for nu in range(2 * lb + 1):
            for mi, m in enumerate(ms):
                terms = []
                for ia in range(na):
                    if Sa[mu][ia] == 0.0:
                        continue
                    for ib in range(nb):
                        if Sb[nu][ib] == 0.0:
                            continue
                        terms.append(f"{fmt(Sa[mu][ia] * Sb[nu][ib])} * acc[{(ia * nb + ib) * len(ms) + mi}]")
                w(f"    out[(row0 + {mu * (2 * lb + 1) + nu}) * ldo + col + {m}] = {' + '.join(terms)};")
    return lines


def gen_class(la, lb, lc, real="double"):
    L = la + lb + lc
    groups = mgroups(la, lb, lc, ACC_LIMIT32 if real == "float" else ACC_LIMIT)
    na, nb = len(cart(la)), len(cart(lb))
    rs = 4 + len(record_layout(la, lb))
    f32 = real == "float"
    lit = lambda x: fmtr(x, real)
    out = []
    w = out.append
    name = f"int3c2e_gen{'32' if f32 else ''}_{la}{lb}{lc}"
    w(f"// class ({la}{lb}|{lc}), {'FP32' if f32 else 'FP64'}: L = {L}, record {rs} values, {len(groups)} m-group(s)")
    mw = None if f32 else MIN_WAVES.get((la, lb, lc))
    bounds = f"64, {mw}" if mw else "64"
    w(f"__global__ __launch_bounds__({bounds}) void {name}(const double* __restrict__ recs, const float* __restrict__ recs32,")
    w(f"    const int64_t* __restrict__ recoff, const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa,")
    w(f"    const double* __restrict__ boys, double* __restrict__ out, int64_t ldo) {{")
    # blockIdx.x = 8 (G (pidx / 8) + group) + pidx % 8: the m-groups of one
    # (pair, aux chunk) block get workgroup ids congruent mod 8 and adjacent in
    # that residue's sequence, so they run side by side on one XCD (workgroups
    # go to the 8 XCDs round robin) and their interleaved column stores into
    # the same lines merge in that XCD's L2 instead of reaching HBM as partial
    # lines, one per group
    w(f"  const int xs = (int)(blockIdx.x >> 3), grp = xs % {len(groups)};")
    w(f"  const int pidx = (xs / {len(groups)}) * 8 + (int)(blockIdx.x & 7);")
    w(f"  if (pidx >= pa.npairs) return;")
    w(f"  const int cend = min(xa.hi, pa.nsurv[pidx]);")
    w(f"  const int cbeg = pa.nbeg ? pa.nbeg[pidx] : 0;")
    w(f"  const int c0 = xa.lo + (int)blockIdx.y * 64, c = c0 + (int)threadIdx.x;")
    w(f"  if (c0 >= cend || c0 + 64 <= cbeg) return;")
    w(f"  const bool active = c >= cbeg && c < cend;")
    w(f"  const int sh = active ? xa.shell[c] : xa.shell[xa.lo];")
    w(f"  const double Cx = ket.xyz[3 * sh], Cy = ket.xyz[3 * sh + 1], Cz = ket.xyz[3 * sh + 2];")
    w(f"  const bool live = active && triplet_kept(pa, xa, pidx, sh, {lc}, Cx, Cy, Cz);")
    w(f"  const int nprim_c = live ? ket.nprim[sh] : 0, pc0 = ket.pptr[sh];")
    w(f"  const int npp = npp_v[pidx];")
    w(f"  const double* __restrict__ recp = recs + recoff[pidx];")
    if f32:
        w(f"  const float* __restrict__ recp32 = recs32 + recoff[pidx];")
    w(f"  const int64_t row0 = pa.row[pidx];")
    w(f"  const int64_t col = active ? xa.col[c] : 0;")
    w(f"  switch (grp) {{")
    for gi, ms in enumerate(groups):
        nacc = na * nb * len(ms)
        w(f"  case {gi}: {{")
        w(f"    double acc[{nacc}];")
        w(f"#pragma unroll")
        w(f"    for (int q = 0; q < {nacc}; q++) acc[q] = 0.0;")
        w(f"    for (int ij = 0; ij < npp; ij++) {{")
        w(f"      const double* __restrict__ rec = recp + (size_t)ij * {rs};")
        w(f"      const double p = rec[0];")
        if f32:
            w(f"      const float* __restrict__ rf = recp32 + (size_t)ij * {rs};")
            w(f"      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;")
            w(f"      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;")
            w(f"      float accf[{nacc}];")
            w(f"#pragma unroll")
            w(f"      for (int q = 0; q < {nacc}; q++) accf[q] = 0.0f;")
        else:
            w(f"      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;")
        w(f"      for (int kp = 0; kp < nprim_c; kp++) {{")
        w(f"      const double gam = ket.exps[pc0 + kp];")
        w(f"      const double alpha = p * gam / (p + gam);")
        if f32:
            w(f"      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);")
        else:
            w(f"      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);")
        w(f"      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];")
        if lc > 0:
            w(f"      {{ const double i2g = 0.5 / gam; sc *= {'-' if lc % 2 else ''}{' * '.join(['i2g'] * lc)}; }}")
        w(f"      double F[{L + 1}];")
        w(f"      boys_array<{L}>(T, boys, F);")
        w(f"      const double ma = -alpha;")
        cvt = "(float)" if f32 else ""
        w(f"      const {real} g0 = {cvt}(sc * F[0]);")
        for n in range(1, L + 1):
            w(f"      sc *= ma; const {real} g{n} = {cvt}(sc * F[{n}]);")
        for d, X in (("x", "PCx"), ("y", "PCy"), ("z", "PCz")):
            w(f"      const {real} p{d}0 = {lit(1.0)}, p{d}1 = {lit(2.0)} * {X};")
            for a in range(2, L + 1):
                w(f"      const {real} p{d}{a} = p{d}{a - 1} * p{d}1;")
        out.extend(gen_group(la, lb, lc, ms, real))
        w(f"      }}")  # kp
        if f32:
            w(f"#pragma unroll")
            w(f"      for (int q = 0; q < {nacc}; q++) acc[q] += (double)accf[q];")
        w(f"    }}")  # ij
        w(f"    if (active) {{")
        out.extend(gen_store(la, lb, lc, ms))
        w(f"    }}")
        w(f"    break;")
        w(f"  }}")
    w(f"  default: break;")
    w(f"  }}")
    w(f"}}")
    w("")
    return out, len(groups), rs


def main(outdir):
    outdir = Path(outdir)
    decls = []
    for la in range(LA_MAX + 1):
        for lb in range(la + 1):
            src = [f"// Generated by gen3c.py -- 3c2e classes ({la}{lb}|lc), lc <= {LC_MAX}.",
                   '#include "gen3c_common.h"', "", "namespace dfhip {", ""]
            regs = []
            for lc in range(LC_MAX + 1):
                body, G, rs = gen_class(la, lb, lc)
                src.extend(body)
                body32, G32, _ = gen_class(la, lb, lc, "float")
                src.extend(body32)
                regs.append(f"  t[{la}][{lb}][{lc}] = Gen3Launch{{&int3c2e_gen_{la}{lb}{lc}, "
                            f"&int3c2e_gen32_{la}{lb}{lc}, {G}, {G32}, {rs}}};")
            src.append(f"void register_gen3_{la}{lb}(Gen3Table& t) {{")
            src.extend(regs)
            src.append("}")
            src.append("")
            src.append("}  // namespace dfhip")
            (outdir / f"int3c2e_gen_{la}{lb}.hip").write_text("\n".join(src) + "\n")
            decls.append(f"void register_gen3_{la}{lb}(Gen3Table& t);")
    hdr = ["// Generated by gen3c.py: registration of the register-resident 3c2e kernels.",
           "#pragma once", '#include "gen3c_common.h"', "", "namespace dfhip {",
           f"static_assert(GEN3_LA_MAX == {LA_MAX} && GEN3_LC_MAX == {LC_MAX}, \"regenerate\");"]
    hdr += decls
    hdr.append("inline void register_gen3_all(Gen3Table& t) {")
    hdr += [f"  register_gen3_{la}{lb}(t);" for la in range(LA_MAX + 1) for lb in range(la + 1)]
    hdr.append("}")
    hdr.append("}  // namespace dfhip")
    (outdir / "gen3c_classes.h").write_text("\n".join(hdr) + "\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else Path(__file__).resolve().parent)
