"""Independent high-precision evaluation of the pair histogram (mpmath, 40 digits).

Derived from textbook geometry rather than from the oracle's binary64 vector formulas, so that a
dropped term, a wrong sign or a transposed operand in the oracle shows up as a mismatch:
  * flat: r = |dx|; 2D: (du_L, du_T) are du rotated into the frame of dx (angle atan2(dy, dx));
    3D: du_L = du.dx / r, transverse energy |du x dx|^2 / r^2, du_T = du . normalize(zhat x dx);
  * sphere: separation by the haversine formula; the longitudinal direction at each endpoint is
    (sin a, cos a) in local (east, north) for the geodesic bearing a (initial bearing at A,
    final bearing at B), the transverse direction is its quarter turn zhat x t = (-cos a, sin a).
Operators follow src/StructureFunctionTypes.jl:43-151.
"""
import mpmath as mp

mp.mp.dps = 40


def op_value(op, L, T, tn2, n2, D):
    kind, nl, nt = op
    if kind == 0:
        ex = mp.mpf(1)
        if nl:
            ex *= L ** nl
        if nt == 2:
            ex *= tn2
        elif nt:
            ex *= T ** nt
        return ex
    if kind == 1:
        return n2
    if kind == 2:
        return L * n2
    if kind == 3:
        return tn2 / (D - 1)
    if kind == 4:
        return L * tn2 / (D - 1)
    if nl == 2:
        return n2
    return mp.sqrt(n2) ** nl


def flat_pair(xi, xj, ui, uj, op):
    D = len(xi)
    dx = [mp.mpf(xj[k]) - mp.mpf(xi[k]) for k in range(D)]
    du = [mp.mpf(uj[k]) - mp.mpf(ui[k]) for k in range(D)]
    r = mp.sqrt(sum(d * d for d in dx))
    n2 = sum(d * d for d in du)
    if D == 2:
        th = mp.atan2(dx[1], dx[0])
        L = du[0] * mp.cos(th) + du[1] * mp.sin(th)
        T = -du[0] * mp.sin(th) + du[1] * mp.cos(th)
        tn2 = T * T
    else:
        L = sum(du[k] * dx[k] for k in range(3)) / r
        c = [du[1] * dx[2] - du[2] * dx[1], du[2] * dx[0] - du[0] * dx[2], du[0] * dx[1] - du[1] * dx[0]]
        tn2 = sum(v * v for v in c) / (r * r)
        h = mp.sqrt(dx[0] ** 2 + dx[1] ** 2)
        T = (du[1] * dx[0] - du[0] * dx[1]) / h
    return r, op_value(op, L, T, tn2, n2, D)


def bearing(f1, l1, f2, l2):
    return mp.atan2(mp.sin(l2 - l1) * mp.cos(f2),
                    mp.cos(f1) * mp.sin(f2) - mp.sin(f1) * mp.cos(f2) * mp.cos(l2 - l1))


def sphere_pair(xi, xj, ui, uj, op, radius=1.0, degrees=True):
    D = len(ui)
    conv = (lambda a: mp.radians(mp.mpf(a))) if degrees else (lambda a: mp.mpf(a))
    l1, f1, l2, f2 = conv(xi[0]), conv(xi[1]), conv(xj[0]), conv(xj[1])
    a = mp.sin((f2 - f1) / 2) ** 2 + mp.cos(f1) * mp.cos(f2) * mp.sin((l2 - l1) / 2) ** 2
    r = 2 * mp.mpf(radius) * mp.asin(mp.sqrt(a))
    aA = bearing(f1, l1, f2, l2)
    aB = bearing(f2, l2, f1, l1) + mp.pi
    tA = (mp.sin(aA), mp.cos(aA))
    tB = (mp.sin(aB), mp.cos(aB))
    nA = (-tA[1], tA[0])
    nB = (-tB[1], tB[0])
    uA = [mp.mpf(v) for v in ui]
    uB = [mp.mpf(v) for v in uj]
    L = uB[0] * tB[0] + uB[1] * tB[1] - (uA[0] * tA[0] + uA[1] * tA[1])
    T = uB[0] * nB[0] + uB[1] * nB[1] - (uA[0] * nA[0] + uA[1] * nA[1])
    tn2 = T * T
    if D == 3:
        tn2 += (uB[2] - uA[2]) ** 2
    n2 = L * L + tn2
    return r, op_value(op, L, T, tn2, n2, D)


def brute(x, u, edges, op, geometry="flat", radius=1.0, vedges=None):
    """Exact-ish histogram: (sums as float, counts, abs sums, min relative distance to an edge)."""
    import numpy as np
    N = x.shape[1]
    nd = len(edges) - 1
    nv = 0 if vedges is None else len(vedges) - 1
    shape = (nd,) if nv == 0 else (nd, nv)
    sums = [[mp.mpf(0) for _ in range(max(nv, 1))] for _ in range(nd)]
    counts = np.zeros(shape, dtype=np.uint64)
    abss = np.zeros(shape)
    margin = mp.inf
    E = [mp.mpf(e) for e in edges]
    VE = None if vedges is None else [mp.mpf(e) for e in vedges]
    for i in range(N):
        for j in range(i + 1, N):
            if geometry == "flat":
                r, v = flat_pair(x[:, i], x[:, j], u[:, i], u[:, j], op)
            else:
                r, v = sphere_pair(x[:, i], x[:, j], u[:, i], u[:, j], op, radius, geometry == "deg")
            for e in E:
                if e != 0:
                    margin = min(margin, abs(r - e) / abs(e))
            b = sum(1 for e in E if e < r)
            if not (1 <= b <= nd):
                continue
            if nv:
                for e in VE:
                    if e != 0:
                        margin = min(margin, abs(v - e) / abs(e))
                    else:
                        margin = min(margin, abs(v))
                vb = sum(1 for e in VE if e < v)
                if not (1 <= vb <= nv):
                    continue
                sums[b - 1][vb - 1] += v
                counts[b - 1, vb - 1] += 1
                abss[b - 1, vb - 1] += float(abs(v))
            else:
                sums[b - 1][0] += v
                counts[b - 1] += 1
                abss[b - 1] += float(abs(v))
    s = np.array([[float(c) for c in row] for row in sums])
    s = s[:, 0] if nv == 0 else s
    return s, counts, abss, float(margin)
