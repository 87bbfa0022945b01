"""The chunked bulge chases equal the unblocked sweep, bit for bit.

The windowed batched QR kernels (z_sweep_win, d_sweep_win) and the large
real path (lchunk_kernel) chase each sweep in chunks and apply the chunk's
reflectors to the rest of H and to Z afterwards, as chains.  These tests
restate the kernels' index logic in numpy, with bounds-checked accesses, and
compare the result with the unblocked sweep of the reference
(GenericSchur.jl:374-504 complex, :837-952 real) in the same floating-point
operations: H and Z must be bitwise equal, and the regions that the large
path's workgroups write in one launch must be disjoint.  CPU only.
"""
import numpy as np
import pytest

class Chk:
    """array wrapper with 0-based bounds-checked access"""
    def __init__(self, a):
        self.a = a

    def __getitem__(self, k):
        for x, m in zip(k, self.a.shape):
            assert 0 <= x < m, (k, self.a.shape)
        return self.a[k]

    def __setitem__(self, k, v):
        for x, m in zip(k, self.a.shape):
            assert 0 <= x < m, (k, self.a.shape)
        self.a[k] = v


def dreflector(nr, a, x1, x2):
    if nr <= 1:
        return 0.0, a, x1, x2
    xnorm = np.hypot(x1, x2) if nr == 3 else abs(x1)
    if xnorm == 0:
        return 0.0, a, x1, x2
    alpha = a
    beta = -np.copysign(np.hypot(alpha, xnorm), alpha)
    tau = (beta - alpha) / beta
    t = 1.0 / (alpha - beta)
    return tau, beta, x1 * t, x2 * t if nr == 3 else x2


def unblocked_sweep(H, Z, mx, istart, iend, v):
    """d_sweep semantics with 1-based indexing helpers on 0-based arrays"""
    n = H.shape[0]
    def HR(r, c):
        return (r - 1, c - 1)
    logs = []
    for k in range(mx, iend):
        nr = min(3, iend - k + 1)
        if k == mx:
            v0, v1, v2s = v
        else:
            v0 = H[HR(k, k - 1)]; v1 = H[HR(k + 1, k - 1)]
            v2s = H[HR(k + 2, k - 1)] if nr == 3 else 0.0
        x1, x2 = v1, (v2s if nr == 3 else 0.0)
        tau1, v0, x1, x2 = dreflector(nr, v0, x1, x2)
        if k > mx:
            H[HR(k, k - 1)] = v0; H[HR(k + 1, k - 1)] = 0.0
            if k < iend - 1:
                H[HR(k + 2, k - 1)] = 0.0
        elif mx > istart:
            H[HR(k, k - 1)] = H[HR(k, k - 1)] * (1.0 - tau1)
        vv2 = x1; tau2 = tau1 * vv2
        if nr == 3:
            vv3 = x2; tau3 = tau1 * vv3
            for j in range(k, n + 1):
                a, b, c = H[HR(k, j)], H[HR(k + 1, j)], H[HR(k + 2, j)]
                ss = a + vv2 * b + vv3 * c
                H[HR(k, j)] = a - ss * tau1; H[HR(k + 1, j)] = b - ss * tau2; H[HR(k + 2, j)] = c - ss * tau3
            for j in range(1, min(k + 3, iend) + 1):
                a, b, c = H[HR(j, k)], H[HR(j, k + 1)], H[HR(j, k + 2)]
                ss = a + vv2 * b + vv3 * c
                H[HR(j, k)] = a - ss * tau1; H[HR(j, k + 1)] = b - ss * tau2; H[HR(j, k + 2)] = c - ss * tau3
        else:
            for j in range(k, n + 1):
                a, b = H[HR(k, j)], H[HR(k + 1, j)]
                ss = a + vv2 * b
                H[HR(k, j)] = a - ss * tau1; H[HR(k + 1, j)] = b - ss * tau2
            for j in range(1, iend + 1):
                a, b = H[HR(j, k)], H[HR(j, k + 1)]
                ss = a + vv2 * b
                H[HR(j, k)] = a - ss * tau1; H[HR(j, k + 1)] = b - ss * tau2
        logs.append((tau1, vv2, x2 if nr == 3 else 0.0))
    # Z chain
    for r in range(1, n + 1):
        for k in range(mx, iend):
            tau1, v2, v3 = logs[k - mx]
            tau2 = tau1 * v2
            if k + 2 <= iend:
                tau3 = tau1 * v3
                a, b, c = Z[r - 1, k - 1], Z[r - 1, k], Z[r - 1, k + 1]
                ss = a + v2 * b + v3 * c
                Z[r - 1, k - 1] = a - ss * tau1; Z[r - 1, k] = b - ss * tau2; Z[r - 1, k + 1] = c - ss * tau3
            else:
                a, b = Z[r - 1, k - 1], Z[r - 1, k]
                ss = a + v2 * b
                Z[r - 1, k - 1] = a - ss * tau1; Z[r - 1, k] = b - ss * tau2


def chain_line(get, put, lg, mx, ka, kb, iend):
    c0 = get(ka); c1 = get(ka + 1)
    k3 = min(kb, iend - 2)
    k = ka
    while k <= k3:
        tau1, v2, v3 = lg[3 * (k - mx)], lg[3 * (k - mx) + 1], lg[3 * (k - mx) + 2]
        tau2 = tau1 * v2; tau3 = tau1 * v3
        b = get(k + 2)
        ss = c0 + v2 * c1 + v3 * b
        put(k, c0 - ss * tau1)
        c0 = c1 - ss * tau2
        c1 = b - ss * tau3
        k += 1
    if k <= kb:
        tau1, v2 = lg[3 * (k - mx)], lg[3 * (k - mx) + 1]
        tau2 = tau1 * v2
        ss = c0 + v2 * c1
        put(k, c0 - ss * tau1)
        c0 = c1 - ss * tau2
        put(k + 1, c0)
        return
    put(kb + 1, c0)
    put(kb + 2, c1)


def chunk_kernel(H, Z, lg, st, ck, wantz, touched, LW):
    n = H.shape[0]
    k0, kend, pk0, pkend, istart, iend, mx = ck
    chase = k0 <= kend
    prev = pk0 <= pkend
    we = min(kend + 3, iend) if chase else 0
    pwe = min(pkend + 3, iend) if prev else 0
    Hc = Chk(H)
    Zc = Chk(Z)
    # WG0
    if chase:
        if prev:
            for lane in range(64):
                col = pwe + 1 + lane
                if col <= we:
                    chain_line(lambda t: Hc[t - 1, col - 1], lambda t, v: (Hc.__setitem__((t - 1, col - 1), v), touched.append(('WG0', t, col))),
                               lg, mx, pk0, pkend, iend)
        wb = max(k0 - 1, 1)
        W = we - wb + 1
        assert W <= LW, (W, LW)
        win = np.zeros((LW, LW))
        for c in range(W):
            for r in range(W):
                win[r, c] = Hc[wb - 1 + r, wb - 1 + c]
        winc = Chk(win)
        def HWg(r, c):
            return winc[r - wb, c - wb]
        def HWs(r, c, v):
            winc[r - wb, c - wb] = v
        for k in range(k0, kend + 1):
            nr = min(3, iend - k + 1)
            if k == mx:
                v0, v1, v2s = st['v']
            else:
                v0 = HWg(k, k - 1); v1 = HWg(k + 1, k - 1)
                v2s = HWg(k + 2, k - 1) if nr == 3 else 0.0
            x1, x2 = v1, (v2s if nr == 3 else 0.0)
            tau1, v0, x1, x2 = dreflector(nr, v0, x1, x2)
            if k > mx:
                HWs(k, k - 1, v0); HWs(k + 1, k - 1, 0.0)
                if k < iend - 1:
                    HWs(k + 2, k - 1, 0.0)
            elif mx > istart:
                HWs(k, k - 1, HWg(k, k - 1) * (1.0 - tau1))
            vv2 = x1; tau2 = tau1 * vv2
            if nr == 3:
                vv3 = x2; tau3 = tau1 * vv3
                for j in range(k, we + 1):
                    p0, p1, p2 = HWg(k, j), HWg(k + 1, j), HWg(k + 2, j)
                    ss = p0 + vv2 * p1 + vv3 * p2
                    HWs(k, j, p0 - ss * tau1); HWs(k + 1, j, p1 - ss * tau2); HWs(k + 2, j, p2 - ss * tau3)
                jm = min(k + 3, iend)
                for j in range(wb, jm + 1):
                    p0, p1, p2 = HWg(j, k), HWg(j, k + 1), HWg(j, k + 2)
                    ss = p0 + vv2 * p1 + vv3 * p2
                    HWs(j, k, p0 - ss * tau1); HWs(j, k + 1, p1 - ss * tau2); HWs(j, k + 2, p2 - ss * tau3)
            else:
                for j in range(k, we + 1):
                    p0, p1 = HWg(k, j), HWg(k + 1, j)
                    ss = p0 + vv2 * p1
                    HWs(k, j, p0 - ss * tau1); HWs(k + 1, j, p1 - ss * tau2)
                for j in range(wb, iend + 1):
                    p0, p1 = HWg(j, k), HWg(j, k + 1)
                    ss = p0 + vv2 * p1
                    HWs(j, k, p0 - ss * tau1); HWs(j, k + 1, p1 - ss * tau2)
            lg[3 * (k - mx)] = tau1; lg[3 * (k - mx) + 1] = vv2; lg[3 * (k - mx) + 2] = x2 if nr == 3 else 0.0
        for c in range(W):
            for r in range(W):
                Hc[wb - 1 + r, wb - 1 + c] = win[r, c]
                touched.append(('WG0', wb + r, wb + c))
    if not prev:
        return
    lcol0 = (we if chase else pwe) + 1
    pwb = max(pk0 - 1, 1)
    nL = (n - lcol0 + 64) // 64 if n >= lcol0 else 0
    nR = (pwb - 1 + 63) // 64
    nZ = (n + 63) // 64 if wantz else 0
    for t in range(nL + nR + nZ):
        for lane in range(64):
            if t < nL:
                col = lcol0 + 64 * t + lane
                if col <= n:
                    chain_line(lambda tt: Hc[tt - 1, col - 1], lambda tt, v: (Hc.__setitem__((tt - 1, col - 1), v), touched.append(('L', tt, col))),
                               lg, mx, pk0, pkend, iend)
            elif t < nL + nR:
                row = 1 + 64 * (t - nL) + lane
                if row <= pwb - 1:
                    chain_line(lambda tt: Hc[row - 1, tt - 1], lambda tt, v: (Hc.__setitem__((row - 1, tt - 1), v), touched.append(('R', row, tt))),
                               lg, mx, pk0, pkend, iend)
            else:
                row = 1 + 64 * (t - nL - nR) + lane
                if row <= n:
                    chain_line(lambda tt: Zc[row - 1, tt - 1], lambda tt, v: Zc.__setitem__((row - 1, tt - 1), v),
                               lg, mx, pk0, pkend, iend)


def chunked_sweep(H, Z, mx, istart, iend, v, LW, wantz=True):
    LKC = LW - 4
    n = H.shape[0]
    lg = np.zeros(3 * n + 3)
    st = {'v': v}
    pk0, pkend = 1, 0
    k0 = mx
    while k0 <= iend - 1:
        kend = min(k0 + LKC - 1, iend - 1)
        touched = []
        chunk_kernel(H, Z, lg, st, (k0, kend, pk0, pkend, istart, iend, mx), wantz, touched, LW)
        # check disjointness of writers in this launch
        owners = {}
        for who, r, c in touched:
            key = (r, c)
            if key in owners and owners[key] != who:
                raise AssertionError(f"overlap {key}: {owners[key]} vs {who}")
            owners[key] = who
        pk0, pkend = k0, kend
        k0 += LKC
    touched = []
    chunk_kernel(H, Z, lg, st, (1, 0, pk0, pkend, istart, iend, mx), wantz, touched, LW)


def zrefl(v0, v1):
    # zreflector2 (householder.jl complex, n=2): returns tau, beta, v1 scaled
    alpha = v0
    xnorm = abs(v1)
    if xnorm == 0 and alpha.imag == 0:
        return 0j, v0, v1
    beta = -np.copysign(np.sqrt(alpha.real ** 2 + alpha.imag ** 2 + xnorm ** 2), alpha.real)
    tau = complex((beta - alpha.real) / beta, -alpha.imag / beta)
    v1 = v1 / (alpha - beta)
    return tau, complex(beta), v1


def sweep_seq(H, Z, istart1, istart, iend, v0, v1):
    n = H.shape[0]
    h = lambda r, c: (r - 1, c - 1)
    for k in range(istart1, iend):
        if k > istart1:
            v0 = H[h(k, k - 1)]; v1 = H[h(k + 1, k - 1)]
        tau1, v0, v1 = zrefl(v0, v1)
        if k > istart1:
            H[h(k, k - 1)] = v0; H[h(k + 1, k - 1)] = 0
        v2 = v1
        tau2 = (tau1 * v2).real
        ct1 = np.conj(tau1)
        for j in range(k, n + 1):
            a, b = H[h(k, j)], H[h(k + 1, j)]
            ss = ct1 * a + tau2 * b
            H[h(k, j)] = a - ss; H[h(k + 1, j)] = b - ss * v2
        jm = min(k + 2, iend)
        for j in range(1, jm + 1):
            a, b = H[h(j, k)], H[h(j, k + 1)]
            ss = tau1 * a + tau2 * b
            H[h(j, k)] = a - ss; H[h(j, k + 1)] = b - ss * np.conj(v2)
        for r in range(1, n + 1):
            a, b = Z[r - 1, k - 1], Z[r - 1, k]
            ss = tau1 * a + tau2 * b
            Z[r - 1, k - 1] = a - ss; Z[r - 1, k] = b - ss * np.conj(v2)
        if k == istart1 and istart1 > istart:
            t = 1 - tau1
            t = t / abs(t)
            ctt = np.conj(t)
            H[h(istart1 + 1, istart1)] *= ctt
            if istart1 + 2 <= iend:
                H[h(istart1 + 2, istart1 + 1)] *= t
            for c in range(1, n + 1):
                cin = istart1 <= c <= iend and c != istart1 + 1
                for r in range(1, min(c - 1, iend) + 1):
                    rin = r >= istart1 and r != istart1 + 1
                    if not rin and not cin:
                        continue
                    x = H[h(r, c)]
                    if rin:
                        x = x * t
                    if cin:
                        x = x * ctt
                    H[h(r, c)] = x
            for c in range(istart1, iend + 1):
                if c != istart1 + 1:
                    Z[:, c - 1] *= ctt
    t = H[h(iend, iend - 1)]
    if t.imag != 0:
        rtv = abs(t); t = t / rtv
        H[h(iend, iend - 1)] = rtv
        for c in range(iend + 1, n + 1):
            H[h(iend, c)] *= np.conj(t)
        for r in range(1, iend):
            H[h(r, iend)] *= t
        Z[:, iend - 1] *= t


def zchain(get, put, lg, ka, kb, left):
    c0 = get(ka)
    for k in range(ka, kb + 1):
        tau1, v2 = lg[2 * (k - ka)], lg[2 * (k - ka) + 1]
        tau2 = (tau1 * v2).real
        b = get(k + 1)
        if left:
            ss = np.conj(tau1) * c0 + tau2 * b
            put(k, c0 - ss); c0 = b - ss * v2
        else:
            ss = tau1 * c0 + tau2 * b
            put(k, c0 - ss); c0 = b - ss * np.conj(v2)
    put(kb + 1, c0)


def sweep_win(H, Z, istart1, istart, iend, v0, v1, ZW):
    ZKC = ZW - 3
    n = H.shape[0]
    h = lambda r, c: (r - 1, c - 1)
    phase_pending = istart1 > istart
    k0 = istart1
    while k0 <= iend - 1:
        kend = k0 if phase_pending else min(k0 + ZKC - 1, iend - 1)
        wb = max(k0 - 1, 1); we = min(kend + 2, iend)
        W = we - wb + 1
        assert W <= ZW
        win = H[wb - 1:we, wb - 1:we].copy()
        WI = lambda r, c: (r - wb, c - wb)
        lg = []
        nv0 = nv1 = 0
        for k in range(k0, kend + 1):
            if k > istart1:
                if k > k0:
                    v0, v1 = nv0, nv1
                else:
                    v0 = win[WI(k, k - 1)]; v1 = win[WI(k + 1, k - 1)]
            tau1, v0, v1 = zrefl(v0, v1)
            if k > istart1:
                win[WI(k, k - 1)] = v0; win[WI(k + 1, k - 1)] = 0
            v2 = v1
            tau2 = (tau1 * v2).real
            ct1 = np.conj(tau1)
            for j in range(k, we + 1):
                a, b = win[WI(k, j)], win[WI(k + 1, j)]
                ss = ct1 * a + tau2 * b
                win[WI(k, j)] = a - ss; win[WI(k + 1, j)] = b - ss * v2
            jm = min(k + 2, iend)
            na = {}
            for j in range(wb, jm + 1):
                a, b = win[WI(j, k)], win[WI(j, k + 1)]
                ss = tau1 * a + tau2 * b
                win[WI(j, k)] = a - ss; win[WI(j, k + 1)] = b - ss * np.conj(v2)
                na[j] = a - ss
            nv0 = na.get(k + 1, 0); nv1 = na.get(k + 2, 0)
            lg += [tau1, v2]
        H[wb - 1:we, wb - 1:we] = win
        for col in range(we + 1, n + 1):
            zchain(lambda t: H[t - 1, col - 1], lambda t, v: H.__setitem__((t - 1, col - 1), v), lg, k0, kend, True)
        for row in range(1, wb):
            zchain(lambda t: H[row - 1, t - 1], lambda t, v: H.__setitem__((row - 1, t - 1), v), lg, k0, kend, False)
        for row in range(1, n + 1):
            zchain(lambda t: Z[row - 1, t - 1], lambda t, v: Z.__setitem__((row - 1, t - 1), v), lg, k0, kend, False)
        if phase_pending:
            phase_pending = False
            t = 1 - lg[0]; t = t / abs(t); ctt = np.conj(t)
            H[h(istart1 + 1, istart1)] *= ctt
            if istart1 + 2 <= iend:
                H[h(istart1 + 2, istart1 + 1)] *= t
            for c in range(1, n + 1):
                cin = istart1 <= c <= iend and c != istart1 + 1
                for r in range(1, min(c - 1, iend) + 1):
                    rin = r >= istart1 and r != istart1 + 1
                    if not rin and not cin:
                        continue
                    x = H[h(r, c)]
                    if rin:
                        x = x * t
                    if cin:
                        x = x * ctt
                    H[h(r, c)] = x
            for c in range(istart1, iend + 1):
                if c != istart1 + 1:
                    Z[:, c - 1] *= ctt
        k0 = kend + 1
    t = H[h(iend, iend - 1)]
    if t.imag != 0:
        rtv = abs(t); t = t / rtv
        H[h(iend, iend - 1)] = rtv
        for c in range(iend + 1, n + 1):
            H[h(iend, c)] *= np.conj(t)
        for r in range(1, iend):
            H[h(r, iend)] *= t
        Z[:, iend - 1] *= t


def windowed_real_sweep(H, Z, mx, istart, iend, v, DW):
    """d_sweep_win: chunks of DW-4 steps chased in a DW x DW window, then the
    chunk's chains (columns right of the window, rows above it, Z)."""
    n = H.shape[0]
    DKC = DW - 4
    Hc, Zc = Chk(H), Chk(Z)
    k0 = mx
    v0, v1, v2s = v
    while k0 <= iend - 1:
        kend = min(k0 + DKC - 1, iend - 1)
        wb, we = max(k0 - 1, 1), min(kend + 3, iend)
        W = we - wb + 1
        assert W <= DW
        win = Chk(np.zeros((DW, DW)))
        for c in range(W):
            for r in range(W):
                win[r, c] = Hc[wb - 1 + r, wb - 1 + c]
        g = lambda r, c: win[r - wb, c - wb]
        def s(r, c, x):
            win[r - wb, c - wb] = x
        lg = np.zeros(3 * DKC)
        for k in range(k0, kend + 1):
            nr = min(3, iend - k + 1)
            if k > mx:
                v0, v1 = g(k, k - 1), g(k + 1, k - 1)
                v2s = g(k + 2, k - 1) if nr == 3 else 0.0
            x1, x2 = v1, (v2s if nr == 3 else 0.0)
            tau1, v0, x1, x2 = dreflector(nr, v0, x1, x2)
            if k > mx:
                s(k, k - 1, v0); s(k + 1, k - 1, 0.0)
                if k < iend - 1:
                    s(k + 2, k - 1, 0.0)
            elif mx > istart:
                s(k, k - 1, g(k, k - 1) * (1.0 - tau1))
            vv2, tau2 = x1, tau1 * x1
            vv3 = x2 if nr == 3 else 0.0
            tau3 = tau1 * vv3
            for j in range(k, we + 1):
                if nr == 3:
                    p0, p1, p2 = g(k, j), g(k + 1, j), g(k + 2, j)
                    ss = p0 + vv2 * p1 + vv3 * p2
                    s(k, j, p0 - ss * tau1); s(k + 1, j, p1 - ss * tau2); s(k + 2, j, p2 - ss * tau3)
                else:
                    p0, p1 = g(k, j), g(k + 1, j)
                    ss = p0 + vv2 * p1
                    s(k, j, p0 - ss * tau1); s(k + 1, j, p1 - ss * tau2)
            jm = min(k + 3, iend) if nr == 3 else iend
            for j in range(wb, jm + 1):
                if nr == 3:
                    p0, p1, p2 = g(j, k), g(j, k + 1), g(j, k + 2)
                    ss = p0 + vv2 * p1 + vv3 * p2
                    s(j, k, p0 - ss * tau1); s(j, k + 1, p1 - ss * tau2); s(j, k + 2, p2 - ss * tau3)
                else:
                    p0, p1 = g(j, k), g(j, k + 1)
                    ss = p0 + vv2 * p1
                    s(j, k, p0 - ss * tau1); s(j, k + 1, p1 - ss * tau2)
            lg[3 * (k - k0):3 * (k - k0) + 3] = (tau1, vv2, vv3)
        for c in range(W):
            for r in range(W):
                Hc[wb - 1 + r, wb - 1 + c] = win[r, c]
        for col in range(we + 1, n + 1):
            chain_line(lambda t: Hc[t - 1, col - 1], lambda t, x: Hc.__setitem__((t - 1, col - 1), x), lg, k0, k0,
                       kend, iend)
        for row in range(1, wb):
            chain_line(lambda t: Hc[row - 1, t - 1], lambda t, x: Hc.__setitem__((row - 1, t - 1), x), lg, k0, k0,
                       kend, iend)
        for row in range(1, n + 1):
            chain_line(lambda t: Zc[row - 1, t - 1], lambda t, x: Zc.__setitem__((row - 1, t - 1), x), lg, k0, k0,
                       kend, iend)
        k0 = kend + 1


def _hessenberg(A):
    import scipy.linalg as sl
    H, Q = sl.hessenberg(A, calc_q=True)
    return np.triu(H, -1), Q


REAL_CASES = [(1, 1, 0), (5, 3, 2), (2, 1, 7), (-30, -35, 0), (-3, -3, 0), (-2, -2, 0), (10, 10, 25)]


@pytest.mark.parametrize("LW", [12, 16, 64])
@pytest.mark.parametrize("case", range(len(REAL_CASES)))
def test_real_large_path_chunks(LW, case):
    """lchunk_kernel: chase in workgroup 0, previous chunk's chains in the
    others, concurrently; bitwise equal to the unblocked double-shift sweep."""
    n = 70
    rng = np.random.default_rng(100 + case)
    H, Z = _hessenberg(rng.standard_normal((n, n)))
    mx, istart, iend = _real_case(REAL_CASES[case], n)
    v = tuple(rng.standard_normal(3))
    H1, Z1, H2, Z2 = H.copy(), Z.copy(), H.copy(), Z.copy()
    unblocked_sweep(H1, Z1, mx, istart, iend, v)
    chunked_sweep(H2, Z2, mx, istart, iend, v, LW)
    assert np.array_equal(H1, H2) and np.array_equal(Z1, Z2)


@pytest.mark.parametrize("DW", [8, 12, 32])
@pytest.mark.parametrize("case", range(len(REAL_CASES)))
def test_real_windowed_sweep(DW, case):
    """d_sweep_win: chunk chased in the window, then its chains, in sequence."""
    n = 70
    rng = np.random.default_rng(200 + case)
    H, Z = _hessenberg(rng.standard_normal((n, n)))
    mx, istart, iend = _real_case(REAL_CASES[case], n)
    v = tuple(rng.standard_normal(3))
    H1, Z1, H2, Z2 = H.copy(), Z.copy(), H.copy(), Z.copy()
    unblocked_sweep(H1, Z1, mx, istart, iend, v)
    windowed_real_sweep(H2, Z2, mx, istart, iend, v, DW)
    assert np.array_equal(H1, H2) and np.array_equal(Z1, Z2)


COMPLEX_CASES = [(1, 1, 0), (5, 3, 0), (3, 1, 10), (20, 20, 5), (7, 2, 1), (-2, -3, 0)]


@pytest.mark.parametrize("ZW", [6, 8, 32])
@pytest.mark.parametrize("case", range(len(COMPLEX_CASES)))
def test_complex_windowed_sweep(ZW, case):
    """z_sweep_win, including the phase scaling after the first step (when
    the sweep starts below istart) and the tail realness scaling."""
    n = 50
    rng = np.random.default_rng(300 + case)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    H, Z = _hessenberg(A)
    c = COMPLEX_CASES[case]
    istart1 = c[0] if c[0] > 0 else n + c[0]
    istart = c[1] if c[1] > 0 else n + c[1]
    iend = n - c[2]
    v0 = complex(rng.standard_normal(), rng.standard_normal())
    v1 = complex(abs(rng.standard_normal()), 0.0)
    H1, Z1, H2, Z2 = H.copy(), Z.copy(), H.copy(), Z.copy()
    sweep_seq(H1, Z1, istart1, istart, iend, v0, v1)
    sweep_win(H2, Z2, istart1, istart, iend, v0, v1, ZW)
    assert np.array_equal(H1, H2) and np.array_equal(Z1, Z2)


def _real_case(c, n):
    mx = c[0] if c[0] > 0 else n + c[0]
    istart = c[1] if c[1] > 0 else n + c[1]
    iend = n - c[2] if c[2] >= 0 else n + c[2]
    return mx, istart, iend


def sweep_win2(H, Z, istart1, istart, iend, v0, v1, ZW):
    """z_sweep_win2: wavefront 0 chases chunk c (after the previous chunk's
    left update of the columns its window adds) while wavefront 1 applies the
    previous chunk's chains; the writes of the two must be disjoint."""
    n = H.shape[0]
    ZKC = ZW - 3
    h = lambda r, c: (r - 1, c - 1)
    phase_pending = istart1 > istart
    pk0, pkend, pwe, pwb = 1, 0, 0, 1
    plg = None
    k0 = istart1
    while True:
        chase = k0 <= iend - 1
        prev = pk0 <= pkend
        if not chase and not prev:
            break
        w0, w1 = set(), set()
        if chase:
            kend = k0 if phase_pending else min(k0 + ZKC - 1, iend - 1)
            wb, we = max(k0 - 1, 1), min(kend + 2, iend)
            if prev:
                for col in range(pwe + 1, we + 1):
                    zchain(lambda t: H[t - 1, col - 1],
                           lambda t, v: (H.__setitem__((t - 1, col - 1), v), w0.add((t, col))), plg, pk0, pkend, True)
            W = we - wb + 1
            assert W <= ZW
            win = H[wb - 1:we, wb - 1:we].copy()
            WI = lambda r, c: (r - wb, c - wb)
            lg = []
            nv0 = nv1 = 0
            for k in range(k0, kend + 1):
                if k > istart1:
                    if k > k0:
                        v0, v1 = nv0, nv1
                    else:
                        v0, v1 = win[WI(k, k - 1)], win[WI(k + 1, k - 1)]
                tau1, v0, v1 = zrefl(v0, v1)
                if k > istart1:
                    win[WI(k, k - 1)] = v0
                    win[WI(k + 1, k - 1)] = 0
                v2 = v1
                tau2 = (tau1 * v2).real
                for j in range(k, we + 1):
                    a, b = win[WI(k, j)], win[WI(k + 1, j)]
                    ss = np.conj(tau1) * a + tau2 * b
                    win[WI(k, j)] = a - ss
                    win[WI(k + 1, j)] = b - ss * v2
                na = {}
                for j in range(wb, min(k + 2, iend) + 1):
                    a, b = win[WI(j, k)], win[WI(j, k + 1)]
                    ss = tau1 * a + tau2 * b
                    win[WI(j, k)] = a - ss
                    win[WI(j, k + 1)] = b - ss * np.conj(v2)
                    na[j] = a - ss
                nv0, nv1 = na.get(k + 1, 0), na.get(k + 2, 0)
                lg += [tau1, v2]
            H[wb - 1:we, wb - 1:we] = win
            w0.update((r, c) for r in range(wb, we + 1) for c in range(wb, we + 1))
        if prev:
            lcol0 = (we if chase else pwe) + 1
            for col in range(lcol0, n + 1):
                zchain(lambda t: H[t - 1, col - 1],
                       lambda t, v: (H.__setitem__((t - 1, col - 1), v), w1.add((t, col))), plg, pk0, pkend, True)
            for row in range(1, pwb):
                zchain(lambda t: H[row - 1, t - 1],
                       lambda t, v: (H.__setitem__((row - 1, t - 1), v), w1.add((row, t))), plg, pk0, pkend, False)
            for row in range(1, n + 1):
                zchain(lambda t: Z[row - 1, t - 1], lambda t, v: Z.__setitem__((row - 1, t - 1), v), plg, pk0, pkend,
                       False)
        assert not (w0 & w1), sorted(w0 & w1)[:5]
        if not chase:
            break
        if phase_pending:
            phase_pending = False
            for col in range(we + 1, n + 1):
                zchain(lambda t: H[t - 1, col - 1], lambda t, v: H.__setitem__((t - 1, col - 1), v), lg, k0, kend, True)
            for row in range(1, wb):
                zchain(lambda t: H[row - 1, t - 1], lambda t, v: H.__setitem__((row - 1, t - 1), v), lg, k0, kend,
                       False)
            for row in range(1, n + 1):
                zchain(lambda t: Z[row - 1, t - 1], lambda t, v: Z.__setitem__((row - 1, t - 1), v), lg, k0, kend,
                       False)
            t = 1 - lg[0]
            t = t / abs(t)
            ctt = np.conj(t)
            H[h(istart1 + 1, istart1)] *= ctt
            if istart1 + 2 <= iend:
                H[h(istart1 + 2, istart1 + 1)] *= t
            for c in range(1, n + 1):
                cin = istart1 <= c <= iend and c != istart1 + 1
                for r in range(1, min(c - 1, iend) + 1):
                    rin = r >= istart1 and r != istart1 + 1
                    if not rin and not cin:
                        continue
                    x = H[h(r, c)]
                    if rin:
                        x = x * t
                    if cin:
                        x = x * ctt
                    H[h(r, c)] = x
            for c in range(istart1, iend + 1):
                if c != istart1 + 1:
                    Z[:, c - 1] *= ctt
            pk0, pkend = 1, 0
        else:
            pk0, pkend, pwe, pwb, plg = k0, kend, we, wb, lg
        k0 = kend + 1
    t = H[h(iend, iend - 1)]
    if t.imag != 0:
        rtv = abs(t)
        t = t / rtv
        H[h(iend, iend - 1)] = rtv
        for c in range(iend + 1, n + 1):
            H[h(iend, c)] *= np.conj(t)
        for r in range(1, iend):
            H[h(r, iend)] *= t
        Z[:, iend - 1] *= t


@pytest.mark.parametrize("ZW", [6, 8, 32])
@pytest.mark.parametrize("case", range(len(COMPLEX_CASES)))
def test_complex_two_wave_sweep(ZW, case):
    """z_sweep_win2: chase of chunk c overlapped with the chains of chunk c-1."""
    n = 50
    rng = np.random.default_rng(400 + case)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    H, Z = _hessenberg(A)
    c = COMPLEX_CASES[case]
    istart1 = c[0] if c[0] > 0 else n + c[0]
    istart = c[1] if c[1] > 0 else n + c[1]
    iend = n - c[2]
    v0 = complex(rng.standard_normal(), rng.standard_normal())
    v1 = complex(abs(rng.standard_normal()), 0.0)
    H1, Z1, H2, Z2 = H.copy(), Z.copy(), H.copy(), Z.copy()
    sweep_seq(H1, Z1, istart1, istart, iend, v0, v1)
    sweep_win2(H2, Z2, istart1, istart, iend, v0, v1, ZW)
    assert np.array_equal(H1, H2) and np.array_equal(Z1, Z2)
