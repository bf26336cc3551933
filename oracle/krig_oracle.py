"""CPU oracle of prediction kriging to wells (test infrastructure only).

Restates Crava::doPostKriging (reference src/crava.cpp:1736-1765) as
CKrigingAdmin runs it with the separable covariances of the posterior
covariance grids, in numpy with the reference's float32 arithmetic:
  * CovGridSeparated(const FFTGrid&) (src/covgridseparated.cpp:16-56):
    gamma_xy(i, j) = C(i, j, 0); gamma_z(k) = C(0, 0, k) / C(0, 0, k0) with
    k0 the end of the first rise of |C(0, 0, k)|;
  * findTaperRanges / angleSum / performTapering (covgridseparated.cpp:
    223-328) and EstimateRangeX/Y/Z (146-221);
  * EstimateSizeOfBlock / EstimateSizeOfBlock2 / CalcCPUTime
    (src/krigingadmin.cpp:768-987);
  * the block loop of KrigAll (krigingadmin.cpp:169-205), the data search
    FindDataInDataBlock / FindDataInDataBlockLoop (441-575, with CBox of
    src/box.cpp), SetMatrix (592-703, its upper triangle used as
    SymmetricMatrix), SetKrigVector (705-766) and KrigBlock (319-398);
  * the data of KrigingData3D / CBWellPt (src/krigingdata3d.cpp:96-181,
    src/bwellpt.cpp:28-80): values of one cell averaged, then the trend at
    the cell subtracted (SubtractTrends, krigingadmin.cpp:400-439).
The block systems are solved with numpy's Cholesky (the reference calls
LAPACK posv through FLENS, which the reference tree does not hold) and the
cell updates k . x are summed in float64.  The functions cosf, sinf and
powf that the reference evaluates in float come from the C library.
Pinned in tests/test_oracle.py by what kriging fixes whatever the planner
does: the kriged value at a data cell equals the datum, one datum gives
trend + gamma(d) / gamma(0) r, cells beyond the covariance window keep the
trend, and the ranges of a known exponential covariance.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

MISSING = np.float32(-99999.0)
_libm = ctypes.CDLL("libm.so.6")
for _f in ("cosf", "sinf"):
    getattr(_libm, _f).restype = ctypes.c_float
    getattr(_libm, _f).argtypes = [ctypes.c_float]
_libm.powf.restype = ctypes.c_float
_libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]
F = np.float32


def _cosf(x):
    return F(_libm.cosf(float(x)))


def _sinf(x):
    return F(_libm.sinf(float(x)))


class CovSep:
    """Separable covariance of one real grid c (nzp, nyp, >= nxp)."""

    def __init__(self, c, nxp, nyp, nzp):
        self.nxp, self.nyp, self.nzp = nxp, nyp, nzp
        self.xy = np.array(c[0, :nyp, :nxp], dtype=np.float32)   # [j, i]
        col = np.array(c[:nzp, 0, 0], dtype=np.float32)
        maxval = F(abs(col[0]))
        i = 1
        value = F(abs(col[1])) if nzp > 1 else F(0)
        while value > maxval and i < nzp - 1:
            maxval = F(abs(col[i]))
            i += 1
            value = F(abs(col[i]))
        maxval = col[i - 1]
        self.z = (col / maxval).astype(np.float32)

    def g2(self, i1, j1, k1, i2, j2, k2):
        """GetGamma2: the covariance from cell 1 to cell 2 (0 beyond half a
        padded grid)."""
        di, dj, dk = i2 - i1, j2 - j1, k2 - k1
        if abs(di) >= self.nxp // 2 or abs(dj) >= self.nyp // 2 or abs(dk) >= self.nzp // 2:
            return F(0)
        return F(self.xy[dj % self.nyp, di % self.nxp] * self.z[dk % self.nzp])

    def g2_many(self, di, dj, dk):
        """g2 for arrays of offsets (cell 2 minus cell 1)."""
        ok = (np.abs(di) < self.nxp // 2) & (np.abs(dj) < self.nyp // 2) & (np.abs(dk) < self.nzp // 2)
        v = self.xy[dj % self.nyp, di % self.nxp] * self.z[dk % self.nzp]
        return np.where(ok, v, np.float32(0)).astype(np.float32)

    def ranges(self):
        """EstimateRangeX / Y / Z (0 where no value reaches the sill)."""
        sill = F(0.05) * self.xy[0, 0]
        nx2, ny2 = self.nxp // 2, self.nyp // 2
        a = np.abs(self.xy)
        rx = ry = rz = 0
        done = False
        for i in range(nx2 - 1, -1, -1):
            for j in list(range(ny2 - 1, -1, -1)) + list(range(ny2, self.nyp)):
                if a[j, i] >= sill:
                    rx, done = i, True
                    break
            if done:
                break
        done = False
        for j in range(ny2 - 1, -1, -1):
            for i in list(range(nx2 - 1, -1, -1)) + list(range(nx2, self.nxp)):
                if a[j, i] >= sill:
                    ry, done = j, True
                    break
            if done:
                break
        for k in range(0, self.nzp // 2 + 1):
            if abs(self.z[k]) <= F(0.05):
                rz = k - 1
                break
        return rx, ry, rz

    def angle_sum(self, angle):
        cosang, sinang = _cosf(angle), _sinf(angle)
        if abs(cosang) > abs(sinang):
            di, dj = F(1), F(sinang / cosang)
        else:
            dj, di = F(cosang / sinang), F(1)
        maxi, maxj = self.nxp // 2 + 1, self.nyp // 2 + 1
        fi = fj = F(0)
        s = F(0)
        while abs(fi) < maxi and abs(fj) < maxj:
            i = int(math.floor(fi)) if fi >= 0 else self.nxp + int(math.floor(fi))
            j = int(math.floor(fj)) if fj >= 0 else self.nyp + int(math.floor(fj))
            s = F(s + self.xy[j, i])
            fi = F(fi + di)
            fj = F(fj + dj)
        return F(s / self.xy[0, 0])

    def taper_ranges(self):
        s = F(0)
        for k in range(self.nzp // 2):
            s = F(s + abs(self.z[k]))
        rz = F(F(2.5) * s)
        angle = F(0)
        da = F(F(math.pi) / F(20))
        rx = ry = F(0)
        while angle < math.pi:
            s = F(F(2.5) * self.angle_sum(angle))
            if F(s * _cosf(angle)) > rx:
                rx = F(s * _cosf(angle))
            if F(s * _sinf(angle)) > ry:
                ry = F(s * _sinf(angle))
            angle = F(angle + da)
        return rx, ry, rz

    def taper(self, rx, ry, rz):
        sq3 = F(np.sqrt(F(3)))

        def w(n, r):
            x = float(F(F(n) / F(r)) * sq3)
            return F(math.exp(-(x * x)))

        nz2 = int(math.ceil(F(self.nzp) / F(2)))
        for k in range(nz2):
            self.z[k] = F(self.z[k] * w(k, rz))
            self.z[self.nzp - 1 - k] = F(self.z[self.nzp - 1 - k] * w(k + 1, rz))
        nx2 = int(math.ceil(F(self.nxp) / F(2)))
        ny2 = int(math.ceil(F(self.nyp) / F(2)))
        wx = np.array([w(i, rx) for i in range(nx2)], np.float32)
        wx1 = np.array([w(i + 1, rx) for i in range(nx2)], np.float32)
        for i in range(nx2):
            self.xy[:, i] = (self.xy[:, i] * wx[i]).astype(np.float32)
            self.xy[:, self.nxp - i - 1] = (self.xy[:, self.nxp - i - 1] * wx1[i]).astype(np.float32)
        wy = np.array([w(j, ry) for j in range(ny2)], np.float32)
        wy1 = np.array([w(j + 1, ry) for j in range(ny2)], np.float32)
        for j in range(ny2):
            self.xy[j, :] = (self.xy[j, :] * wy[j]).astype(np.float32)
            self.xy[self.nyp - j - 1, :] = (self.xy[self.nyp - j - 1, :] * wy1[j]).astype(np.float32)


def _box(a0, b0, c0, a1, b1, c1, dims=None):
    b = [min(a0, a1), min(b0, b1), min(c0, c1), max(a0, a1), max(b0, b1), max(c0, c1)]
    if dims is not None:
        for q in range(3):
            b[q] = min(max(b[q], 0), dims[q] - 1)
            b[q + 3] = min(max(b[q + 3], 0), dims[q] - 1)
    return b


def _half(b, s):
    """CBox::ModifyBox(s) without ranges: halfway (float halves truncated)."""
    return [int(F(b[q] + s[q]) / F(2)) for q in range(6)]


class Planner:
    """The block structure and systems of CKrigingAdmin for one data set."""

    def __init__(self, covs, dims, pts, target):
        self.cov = covs
        self.nx, self.ny, self.nz = dims
        self.pts = pts            # list of [i, j, k, values (3 float32), valid (3 bool)]
        self.target = target
        self.nvalid = sum(int(v) for p in pts for v in p[4])
        self._block_size()

    def _cpu_time(self, dxB, dyE):
        rX, rY, rZ = self.rangeX, self.rangeY, self.rangeZ
        dxE = F(F(F(dyE) * rX) / rY)
        dzE = F(F(F(dyE) * rZ) / rY)
        dyB = F(F(F(dxB) * rY) / rX)
        dzB = F(F(F(dxB) * rZ) / rX)
        V = F(self.nx * self.ny * self.nz)
        v = F(F(F(dxB) * dyB) * dzB)
        Nss = F(V / v)
        nd = F(F(F(F(F(F(dxB) + F(2) * dxE) * F(dyB + F(2) * dyE)) * F(dzB + F(2) * dzE)) * F(self.nvalid)) / V)
        TBig = F(F(F(F(1.50e-6) * Nss) * nd) * nd)
        TChol = F(F(F(F(F(1.90e-8) * Nss) * nd) * nd) * nd)
        TSolve = F(F(F(F(7.00e-7) * Nss) * nd) * nd)
        TSmall = F(F(F(1.64e-6) * V) * nd)
        rapid = float(TChol) >= 0.5 * float(TSmall)
        return F(F(F(TBig + TChol) + TSolve) + TSmall), nd, rapid

    def _size2(self):
        rX, rY, rZ = self.rangeX, self.rangeY, self.rangeZ
        self.dyB = int(math.ceil(F(F(self.dxB) * rY) / rX))
        self.dzB = int(math.ceil(F(F(self.dxB) * rZ) / rX))
        if self.dyE > int(rY):
            self.dyE = int(rY)
        self.dxE = int(math.ceil(F(F(self.dyE) * rX) / rY))
        self.dzE = int(math.ceil(F(F(self.dyE) * rZ) / rY))
        self.dxE = min(self.dxE, int(rX))
        self.dzE = min(self.dzE, int(rZ))

    def _block_size(self):
        tr = [self.cov[a].taper_ranges() for a in range(3)]
        rx = max(max(tr[0][0], tr[1][0]), tr[2][0])
        ry = max(max(tr[0][1], tr[1][1]), tr[2][1])
        rz = max(max(tr[0][2], tr[1][2]), tr[2][2])
        for c in self.cov:
            c.taper(rx, ry, rz)
        r = [self.cov[a].ranges() for a in range(3)]
        self.rangeX = F(max(max(r[0][0], r[1][0]), r[2][0]))
        self.rangeY = F(max(max(r[0][1], r[1][1]), r[2][1]))
        self.rangeZ = F(max(max(r[0][2], r[1][2]), r[2][2]))
        self.taper_r = (rx, ry, rz)
        nx, ny, nz = self.nx, self.ny, self.nz
        rX, rY, rZ = self.rangeX, self.rangeY, self.rangeZ
        if len(self.pts) <= self.target:
            self.dxB, self.dyB, self.dzB = nx, ny, nz
            self.dxE, self.dyE, self.dzE = int(rX), int(rY), int(rZ)
        elif rX == 0 and rY == 0:
            self.dxB, self.dyB, self.dxE, self.dyE = 1, 1, 0, 0
            self.dzB = int(math.ceil(rZ))
            self.dzE = self.dzB
        elif rX == 0:
            self.dxB, self.dxE = 1, 0
            self.dyB = self.dyE = int(math.ceil(rY))
            self.dzB = self.dzE = int(math.ceil(rZ))
        elif rY == 0:
            self.dyB, self.dyE = 1, 0
            self.dxB = self.dxE = int(math.ceil(rX))
            self.dzB = self.dzE = int(math.ceil(rZ))
        else:
            dyE = int(rY)
            tMin, nd, rapid = self._cpu_time(1, dyE)
            ndT, dxT = nd, 0
            for dxB in range(2, nx + 1):
                t1, nd, rapid = self._cpu_time(dxB, dyE)
                if t1 < tMin:
                    tMin, dxT, ndT = t1, dxB, nd
            if ndT < self.target:
                self.dxB, self.dyE = dxT, dyE
                self._size2()
            else:
                self._cpu_time(1, 1)
                _, nd, rapid = self._cpu_time(nx, float(rY))
                V = F(nx * ny * nz)
                R = F(F(rX * rY) * rZ)
                a = F(_libm.powf(float(F(F(F(self.target) * V) / F(F(self.nvalid) * R))), float(F(0.333333333))))
                maxdy = int(F(F(F(0.5) * F(a - F(F(1) / rX))) * rY))
                test, sdy, sdx, cnt = False, F(0), F(0), 0
                for dy in range(1, maxdy + 1):
                    dx2 = F(rX * F(a - F(F(2 * dy) / rY)))
                    if dx2 < 1:
                        continue
                    if rapid:
                        test, sdy, sdx, cnt = True, F(sdy + dy), F(sdx + dx2), cnt + 1
                        break
                if test:
                    self.dxB = int(math.ceil(F(sdx / F(cnt))))
                    self.dyE = int(math.ceil(F(sdy / F(cnt))))
                    self._size2()
                else:
                    self.dxB, self.dyB, self.dzB = nx, ny, nz
                    self.dxE, self.dyE, self.dzE = int(rX), int(rY), int(rZ)
        self.dxB, self.dyB, self.dzB = max(self.dxB, 1), max(self.dyB, 1), max(self.dzB, 1)

    def find(self, g, box):
        lists = [[], [], []]
        tot = 0
        cmin = int(F(self.target) * F(F(1) - F(10 / np.float32(100))))
        cmax = int(F(self.target) * F(F(1) + F(10 / np.float32(100))))
        for m, p in enumerate(self.pts):
            i, j, k = p[0], p[1], p[2]
            if not (box[0] <= i <= box[3] and box[1] <= j <= box[4] and box[2] <= k <= box[5]):
                continue
            if p[4][g]:
                lists[g].append(m)
                tot += 1
            else:
                for a in range(3):
                    if a != g and p[4][a]:
                        lists[a].append(m)
                        tot += 1
        fit = "right" if cmin <= tot <= cmax else ("small" if tot < cmin else "big")
        return fit, lists

    def find_loop(self, g, block, data):
        cur, lists = self.find(g, data)
        minB = list(data)
        maxB = _box(block[0] - int(self.rangeX), block[1] - int(self.rangeY), block[2] - int(self.rangeZ),
                    block[3] + int(self.rangeX), block[4] + int(self.rangeY), block[5] + int(self.rangeZ))
        if cur == "small":
            t, lists = self.find(g, maxB)
            if t != "big":
                data, cur = maxB, "right"
        elif cur == "big":
            t, lists = self.find(g, minB)
            if t != "small":
                data, cur = minB, "right"
        while cur != "right":
            if cur == "small":
                minB = list(data)
                data = _half(data, maxB)
            else:
                maxB = list(data)
                data = _half(data, minB)
            if data == maxB or data == minB:
                break
            cur, lists = self.find(g, data)
        return lists


TAB_OF = ((0, 3, 4), (3, 1, 5), (4, 5, 2))
FLIP_OF = ((0, 0, 0), (1, 1, 0), (1, 1, 1))


def krig_post(trend, cov_grids, points, data_target, dims, padded):
    """trend: three real (nz, ny, >= nx) arrays (copies are returned kriged);
    cov_grids: six real (nzp, nyp, >= nxp) arrays (aa bb rr ab ar br);
    points: rows (i, j, k, log Vp, log Vs, log rho) with MISSING; dims =
    (nx, ny, nz), padded = (nxp, nyp, nzp).  Returns (kriged trends,
    statistics)."""
    nxp, nyp, nzp = padded
    covs = [CovSep(c, nxp, nyp, nzp) for c in cov_grids]
    return krig_with(trend, covs, points, data_target, dims)


def krig_with(trend, covs, points, data_target, dims):
    nx, ny, nz = dims
    out = [np.array(t, dtype=np.float32) for t in trend]
    cells = {}
    order = []
    for r in points:
        i, j, k = int(r[0]), int(r[1]), int(r[2])
        v = [F(r[3]), F(r[4]), F(r[5])]
        if all(x == MISSING for x in v):
            continue
        if (i, j, k) not in cells:
            cells[(i, j, k)] = [[F(0)] * 3, [0, 0, 0]]
            order.append((i, j, k))
        for a in range(3):
            if v[a] != MISSING:
                cells[(i, j, k)][0][a] = F(cells[(i, j, k)][0][a] + v[a])
                cells[(i, j, k)][1][a] += 1
    pts = []
    for key in order:
        s, n = cells[key]
        vals = [F(s[a] / F(n[a])) if n[a] else F(0) for a in range(3)]
        pts.append([key[0], key[1], key[2], vals, [n[a] > 0 for a in range(3)]])
    P = Planner(covs, dims, pts, data_target)
    for p in pts:
        for a in range(3):
            t = F(out[a][p[2], p[1], p[0]])
            if p[4][a] and t != MISSING:
                p[3][a] = F(p[3][a] - t)
    stats = dict(block=(P.dxB, P.dyB, P.dzB), block_ext=(P.dxE, P.dyE, P.dzE),
                 range=(int(P.rangeX), int(P.rangeY), int(P.rangeZ)), data_points=len(pts), empty_blocks=[0, 0, 0],
                 max_block_data=[0, 0, 0], cells_updated=[0, 0, 0])
    if not pts:
        return out, stats
    nbx, nby, nbz = -(-nx // P.dxB), -(-ny // P.dyB), -(-nz // P.dzB)
    stats["blocks"] = nbx * nby * nbz
    for g in range(3):
        upd = out[g].copy()
        for bk in range(nbz):
            for bj in range(nby):
                for bi in range(nbx):
                    i1, j1, k1 = bi * P.dxB, bj * P.dyB, bk * P.dzB
                    block = _box(i1, j1, k1, i1 + P.dxB - 1, j1 + P.dyB - 1, k1 + P.dzB - 1, dims)
                    data = _box(i1 - P.dxE, j1 - P.dyE, k1 - P.dzE, i1 + P.dxB + P.dxE - 1,
                                j1 + P.dyB + P.dyE - 1, k1 + P.dzB + P.dzE - 1, dims)
                    lists = P.find_loop(g, block, data)
                    idx = [(m, a) for a in range(3) for m in lists[a]]
                    n = len(idx)
                    if n == 0:
                        stats["empty_blocks"][g] += 1
                        continue
                    stats["max_block_data"][g] = max(stats["max_block_data"][g], n)
                    K = np.zeros((n, n))
                    r = np.zeros(n)
                    for a_, (m, ta) in enumerate(idx):
                        pa = pts[m]
                        r[a_] = pa[3][ta]
                        for b_ in range(a_, n):
                            mb, tb = idx[b_]
                            pb = pts[mb]
                            K[a_, b_] = covs[TAB_OF[ta][tb]].g2(pa[0], pa[1], pa[2], pb[0], pb[1], pb[2])
                    Ks = np.triu(K) + np.triu(K, 1).T
                    L = np.linalg.cholesky(Ks)
                    x = np.linalg.solve(L.T, np.linalg.solve(L, r))
                    ii, jj, kk = np.meshgrid(np.arange(block[0], block[3] + 1), np.arange(block[1], block[4] + 1),
                                             np.arange(block[2], block[5] + 1), indexing="ij")
                    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
                    kv = np.zeros((ii.size, n), np.float64)
                    for c_, (m, ta) in enumerate(idx):
                        p = pts[m]
                        q = TAB_OF[g][ta]
                        sg = -1 if FLIP_OF[g][ta] else 1
                        kv[:, c_] = covs[q].g2_many(sg * (p[0] - ii), sg * (p[1] - jj), sg * (p[2] - kk))
                    add = kv @ x
                    t = out[g][kk, jj, ii]
                    ok = t != MISSING
                    upd[kk[ok], jj[ok], ii[ok]] = (t[ok] + add[ok].astype(np.float32)).astype(np.float32)
                    stats["cells_updated"][g] += int(ok.sum())
        out[g] = upd
    return out, stats
