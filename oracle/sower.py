"""Restatement of AERO-F's SOWER readers in numpy -- test infrastructure
(the oracle for aerofx_sower_read / aerofx_solution_read; the library never
uses it).

BinFileHandler (BinFileHandler.h:167-196): int 1 + double version header,
byte order from the int; offsets count from the end of the header.
GeoSource::getSubDomain (GeoSource.C:246-379) with getRangeInfo
(GeoSource.C:204-240), NodeSet::read (Node.C:7-50), ElemSet::read
(ElemCore.C:135-200), FaceSet::read (FaceCore.C:270-347), Connectivity
(BinFileHandler&) (ConnectivityCore.C:869-882); solution files:
SubDomain::readTagFromFile / readVectorFromFile (SubDomain.C:5609-5740).
"""
import struct

import numpy as np

HEADER = 12


class _Bin:
    def __init__(self, path):
        with open(path, "rb") as f:
            self.b = f.read()
        self.e = "<" if struct.unpack("<i", self.b[:4])[0] == 1 else ">"
        self.version = struct.unpack(self.e + "d", self.b[4:12])[0]
        self.pos = HEADER

    def seek(self, off):
        self.pos = HEADER + off

    def tell(self):
        return self.pos - HEADER

    def get(self, fmt, n=1):
        size = struct.calcsize(fmt) * n
        v = struct.unpack(self.e + fmt * n, self.b[self.pos:self.pos + size])
        self.pos += size
        return v if n > 1 else v[0]

    def arr(self, fmt, n):
        return list(self.get(fmt, n)) if n > 1 else ([self.get(fmt)] if n == 1 else [])


def _ranges(f):
    nr, ns = f.get("i"), f.get("i")
    r = f.arr("i", 2 * nr)
    s = f.arr("i", ns)
    return [(r[2 * k], r[2 * k + 1]) for k in range(nr)] + [(x, x) for x in s]


def read_subdomain(geo, dec, sub):
    d = _Bin(dec)
    nsub = d.get("i")
    assert 0 <= sub < nsub
    d.seek(4 + 8 * sub)
    d.seek(d.get("q"))
    nodes, elems, faces = _ranges(d), _ranges(d), _ranges(d)
    nnb = d.get("i")
    neighb, shared = [], None
    if nnb > 0:
        neighb = d.arr("i", nnb)
        size, ntarget = d.get("i"), d.get("i")
        shared = (d.arr("i", size + 1), d.arr("i", ntarget))
    g = _Bin(geo)
    toc = [g.get("q") for _ in range(3)]
    g.seek(toc[0])
    nclus = g.get("i")
    start = g.get("q")
    X, glob, clus = [], [], []
    for lo, hi in nodes:
        g.seek(start + lo * 28)
        for c in range(lo, hi + 1):
            glob.append(g.get("i"))
            X.append(g.arr("d", 3))
            clus.append(c)
    loc = {c: i for i, c in enumerate(clus)}

    def sect(at, width, parse):
        g.seek(at)
        g.get("i")
        st = g.get("q")
        tstart = g.tell()
        out = []
        for lo, hi in (elems if width == 4 else faces):
            g.seek(tstart + 4 * lo)
            g.seek(st + 4 * g.get("i"))
            for _ in range(lo, hi + 1):
                out.append(parse())
        return out

    def tet():
        typ, _, _ = g.get("i"), g.get("i"), g.get("i")
        assert typ == 5
        return [loc[x] for x in g.arr("i", 4)]

    def tri():
        typ, _ = g.get("i"), g.get("i")
        assert typ == 4
        nd = [loc[x] for x in g.arr("i", 3)]
        code = g.get("i")
        g.get("i")
        return nd + [11 if code == 6 else code]

    T = sect(toc[1], 4, tet)
    Fc = sect(toc[2], 3, tri)
    return dict(X=np.array(X, dtype=np.float64).reshape(-1, 3), tets=np.array(T, dtype=np.int32).reshape(-1, 4),
                faces=np.array([f[:3] for f in Fc], dtype=np.int32).reshape(-1, 3),
                codes=np.array([f[3] for f in Fc], dtype=np.int32), node_glob=np.array(glob, dtype=np.int32),
                node_clus=np.array(clus, dtype=np.int32), num_clus_nodes=nclus, neighbours=neighb, shared=shared,
                version=g.version)


def read_solution(path, step, node_clus=None):
    f = _Bin(path)
    n, neq, nsol = f.get("i", 3)
    assert step < nsol
    rec = 12 + step * (8 + n * neq * 8)
    f.seek(rec)
    tag = f.get("d")
    rows = range(n) if node_clus is None else node_clus
    U = []
    for c in rows:
        f.seek(rec + 8 + c * neq * 8)
        U.append(f.arr("d", neq))
    return np.array(U).reshape(-1, neq), tag
