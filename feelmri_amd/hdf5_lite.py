"""Read-only HDF5 for mesh files: the numeric datasets an XDMF file points
at, without h5py (not installed here).

Written from the HDF5 File Format Specification (version 3.0) for what
meshio, dolfin and dolfinx write with the library's default settings:

* superblock versions 0-3; object headers version 1 and 2 (with
  continuation blocks);
* groups as symbol tables (v1 B-tree of symbol-table nodes plus a local
  heap) or as compact link messages; dense link storage (fractal heap) is
  not read and raises;
* datasets of fixed-point or IEEE floating-point elements, either byte
  order; compact, contiguous, or chunked storage indexed by a v1 B-tree,
  with the deflate and shuffle filters (Fletcher-32 checksums are skipped);
  unwritten chunks read as 0.

``File(path)["/group/dataset"]`` returns a numpy array; ``File.keys(group)``
lists a group.  Any other construct raises ``NotImplementedError`` naming
it.
"""
from __future__ import annotations

import zlib

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF


class _Reader:
    def __init__(self, data: bytes, so: int = 8, sl: int = 8):
        self.d = data
        self.so, self.sl = so, sl

    def u(self, pos: int, n: int) -> int:
        return int.from_bytes(self.d[pos:pos + n], "little")

    def off(self, pos: int) -> int:
        return self.u(pos, self.so)

    def length(self, pos: int) -> int:
        return self.u(pos, self.sl)


class _Dataset:
    def __init__(self):
        self.shape = None
        self.dtype = None
        self.layout = None      # ("compact", bytes) | ("contiguous", addr, size) | ("chunked", btree, chunk_dims)
        self.filters = []       # [(id, client_values)]


class File:
    """An HDF5 file opened for reading (the whole file is read into memory)."""

    def __init__(self, path):
        with open(path, "rb") as f:
            data = f.read()
        # the superblock sits at 0 or after a user block of 512 * 2^k bytes
        base, k = 0, 512
        while data[base:base + 8] != SIGNATURE:
            base, k = k, 2 * k
            if base >= len(data):
                raise ValueError(f"{path}: no HDF5 signature")
        self.path = str(path)
        self._d = data
        self._base = base
        self._parse_superblock(base)

    # -- superblock ----------------------------------------------------------------
    def _parse_superblock(self, p: int) -> None:
        d = self._d
        ver = d[p + 8]
        if ver in (0, 1):
            so, sl = d[p + 13], d[p + 14]
            self.r = _Reader(d, so, sl)
            q = p + 24 + (4 if ver == 1 else 0)
            self._base_addr = self.r.off(q)
            q += 4 * so                                  # base, free space, EOF, driver info
            # root group symbol table entry
            self._root = self._symbol_entry(q)[1]
        elif ver in (2, 3):
            so, sl = d[p + 9], d[p + 10]
            self.r = _Reader(d, so, sl)
            q = p + 12
            self._base_addr = self.r.off(q)
            q += 3 * so                                  # base, superblock extension, EOF
            self._root = self.r.off(q)
        else:
            raise NotImplementedError(f"HDF5 superblock version {ver}")
        if self._base_addr == UNDEF:
            self._base_addr = 0
        self._base_addr += self._base if self._base_addr == 0 else 0

    def _a(self, addr: int) -> int:
        return self._base_addr + addr

    def _symbol_entry(self, q: int):
        """(link name offset, object header address, cache type, scratch) at q."""
        r = self.r
        name_off = r.off(q)
        hdr = r.off(q + r.so)
        cache = r.u(q + 2 * r.so, 4)
        scratch = q + 2 * r.so + 8
        return name_off, hdr, cache, scratch

    # -- object headers --------------------------------------------------------------
    def _messages(self, addr: int):
        """[(type, data_offset, size)] of the object header at addr."""
        d, r = self._d, self.r
        p = self._a(addr)
        out = []
        if d[p:p + 4] == b"OHDR":
            ver, flags = d[p + 4], d[p + 5]
            if ver != 2:
                raise NotImplementedError(f"object header version {ver}")
            q = p + 6
            if flags & 0x20:
                q += 16
            if flags & 0x10:
                q += 4
            nsz = 1 << (flags & 3)
            size0 = r.u(q, nsz)
            q += nsz
            blocks = [(q, size0)]
            tracked = bool(flags & 0x04)
            while blocks:
                start, size = blocks.pop(0)
                q, end = start, start + size
                while q + 4 <= end:
                    mtype, msize = d[q], r.u(q + 1, 2)
                    q += 4 + (2 if tracked else 0)
                    if mtype == 0x10:
                        caddr, clen = r.off(q), r.length(q + r.so)
                        c = self._a(caddr)
                        if d[c:c + 4] != b"OCHK":
                            raise ValueError("bad object header continuation block")
                        blocks.append((c + 4, clen - 8))
                    elif mtype != 0:
                        out.append((mtype, q, msize))
                    q += msize
            return out
        ver = d[p]
        if ver != 1:
            raise NotImplementedError(f"object header version {ver}")
        nmsg = r.u(p + 2, 2)
        size = r.u(p + 8, 4)
        blocks = [(p + 16, size)]
        while blocks and len(out) < nmsg:
            start, size = blocks.pop(0)
            q, end = start, start + size
            while q + 8 <= end:
                mtype, msize = r.u(q, 2), r.u(q + 2, 2)
                q += 8
                if mtype == 0x10:
                    blocks.append((self._a(r.off(q)), r.length(q + r.so)))
                elif mtype != 0:
                    out.append((mtype, q, msize))
                q += msize
        return out

    # -- groups ----------------------------------------------------------------------
    def _links(self, addr: int) -> dict:
        """name -> object header address of the group whose header is at addr."""
        d, r = self._d, self.r
        links = {}
        for mtype, q, _ in self._messages(addr):
            if mtype == 0x11:                                # symbol table
                self._symbol_table(r.off(q), r.off(q + r.so), links)
            elif mtype == 0x06:                              # link (compact storage)
                flags = d[q + 1]
                s = q + 2
                ltype = 0
                if flags & 0x08:
                    ltype = d[s]
                    s += 1
                if flags & 0x04:
                    s += 8
                if flags & 0x10:
                    s += 1
                nlen_sz = 1 << (flags & 3)
                nlen = r.u(s, nlen_sz)
                s += nlen_sz
                name = d[s:s + nlen].decode()
                s += nlen
                if ltype == 0:
                    links[name] = r.off(s)
            elif mtype == 0x02:                              # link info
                heap = r.off(q + 2 + (8 if d[q + 1] & 1 else 0))
                if heap != UNDEF:
                    raise NotImplementedError("dense link storage (fractal heap) is not read")
        return links

    def _symbol_table(self, btree: int, heap: int, links: dict) -> None:
        d, r = self._d, self.r
        h = self._a(heap)
        if d[h:h + 4] != b"HEAP":
            raise ValueError("bad local heap")
        seg = self._a(r.off(h + 8 + 2 * r.sl))

        def name_at(o):
            e = d.index(b"\0", seg + o)
            return d[seg + o:e].decode()

        def walk(node):
            n = self._a(node)
            if d[n:n + 4] == b"SNOD":
                cnt = r.u(n + 6, 2)
                q = n + 8
                for _ in range(cnt):
                    name_off, hdr, _, _ = self._symbol_entry(q)
                    links[name_at(name_off)] = hdr
                    q += 2 * r.so + 24
                return
            if d[n:n + 4] != b"TREE" or d[n + 4] != 0:
                raise ValueError("bad group B-tree node")
            used = r.u(n + 6, 2)
            q = n + 8 + 2 * r.so + r.sl                      # past siblings and key 0
            for _ in range(used):
                walk(r.off(q))
                q += r.so + r.sl
        walk(btree)

    def _resolve(self, path: str) -> int:
        addr = self._root
        for part in [p for p in path.split("/") if p]:
            links = self._links(addr)
            if part not in links:
                raise KeyError(f"{self.path}: no '{part}' in {path}")
            addr = links[part]
        return addr

    def keys(self, group: str = "/") -> list:
        return sorted(self._links(self._resolve(group)))

    # -- datasets --------------------------------------------------------------------
    def _dataset(self, addr: int) -> _Dataset:
        d, r = self._d, self.r
        ds = _Dataset()
        for mtype, q, size in self._messages(addr):
            if mtype == 0x01:
                ver, nd, flags = d[q], d[q + 1], d[q + 2]
                s = q + (8 if ver == 1 else 4)
                ds.shape = tuple(r.length(s + i * r.sl) for i in range(nd))
            elif mtype == 0x03:
                ds.dtype = self._dtype(q)
            elif mtype == 0x08:
                ds.layout = self._layout(q)
            elif mtype == 0x0B:
                ds.filters = self._filters(q)
        if ds.shape is None or ds.dtype is None or ds.layout is None:
            raise ValueError("object is not a dataset")
        return ds

    def _dtype(self, q: int) -> np.dtype:
        d, r = self._d, self.r
        cls = d[q] & 0x0F
        bits = d[q + 1]
        size = r.u(q + 4, 4)
        order = ">" if bits & 1 else "<"
        if cls == 0:
            kind = "i" if bits & 0x08 else "u"
        elif cls == 1:
            if bits & 0x40:
                raise NotImplementedError("VAX floating point")
            kind = "f"
        else:
            raise NotImplementedError(f"HDF5 datatype class {cls}")
        return np.dtype(f"{order}{kind}{size}")

    def _layout(self, q: int):
        d, r = self._d, self.r
        ver = d[q]
        if ver in (1, 2):
            nd, cls = d[q + 1], d[q + 2]
            s = q + 8
            addr = None
            if cls != 0:
                addr = r.off(s)
                s += r.so
            dims = [r.u(s + 4 * i, 4) for i in range(nd)]
            s += 4 * nd
            if cls == 0:
                n = r.u(s, 4)
                return ("compact", d[s + 4:s + 4 + n])
            if cls == 1:
                return ("contiguous", addr, None)
            return ("chunked", addr, dims[:-1] if len(dims) > 1 else dims)
        if ver == 3:
            cls = d[q + 1]
            s = q + 2
            if cls == 0:
                n = r.u(s, 2)
                return ("compact", d[s + 2:s + 2 + n])
            if cls == 1:
                return ("contiguous", r.off(s), r.length(s + r.so))
            if cls == 2:
                nd = d[s]
                addr = r.off(s + 1)
                dims = [r.u(s + 1 + r.so + 4 * i, 4) for i in range(nd)]
                return ("chunked", addr, dims[:-1])         # the last dimension is the element size
        raise NotImplementedError(f"data layout message version {ver} class {d[q + 1]}")

    def _filters(self, q: int):
        d, r = self._d, self.r
        ver, n = d[q], d[q + 1]
        s = q + (8 if ver == 1 else 2)
        out = []
        for _ in range(n):
            fid = r.u(s, 2)
            if ver == 1 or fid >= 256:
                nlen = r.u(s + 2, 2)
                s += 4
            else:
                nlen = 0
                s += 2
            nval = r.u(s + 2, 2)
            s += 4
            if ver == 1:
                nlen = (nlen + 7) // 8 * 8
            s += nlen
            vals = [r.u(s + 4 * i, 4) for i in range(nval)]
            s += 4 * nval
            if ver == 1 and nval % 2:
                s += 4
            out.append((fid, vals))
        return out

    def _unfilter(self, raw: bytes, filters, mask: int, itemsize: int) -> bytes:
        for i in reversed(range(len(filters))):
            if mask & (1 << i):
                continue
            fid = filters[i][0]
            if fid == 1:
                raw = zlib.decompress(raw)
            elif fid == 2:
                a = np.frombuffer(raw, dtype=np.uint8)
                n = a.size // itemsize
                body = a[:n * itemsize].reshape(itemsize, n).T.reshape(-1)
                raw = body.tobytes() + a[n * itemsize:].tobytes()
            elif fid == 3:
                raw = raw[:-4]
            else:
                raise NotImplementedError(f"HDF5 filter {fid}")
        return raw

    def _chunks(self, btree: int, nd: int):
        """(offsets, size, filter mask, address) of every chunk of a v1
        chunk B-tree."""
        d, r = self._d, self.r
        key = 8 + 8 * (nd + 1)
        out = []

        def walk(node):
            n = self._a(node)
            if d[n:n + 4] != b"TREE" or d[n + 4] != 1:
                raise ValueError("bad chunk B-tree node")
            level, used = d[n + 5], r.u(n + 6, 2)
            q = n + 8 + 2 * r.so
            for _ in range(used):
                size, mask = r.u(q, 4), r.u(q + 4, 4)
                offs = [r.u(q + 8 + 8 * i, 8) for i in range(nd)]
                child = r.off(q + key)
                if level == 0:
                    out.append((offs, size, mask, child))
                else:
                    walk(child)
                q += key + r.so
        walk(btree)
        return out

    def __getitem__(self, path: str) -> np.ndarray:
        ds = self._dataset(self._resolve(path))
        shape, dt = ds.shape, ds.dtype
        count = int(np.prod(shape)) if shape else 1
        kind = ds.layout[0]
        if kind == "compact":
            arr = np.frombuffer(ds.layout[1], dtype=dt, count=count)
        elif kind == "contiguous":
            addr = ds.layout[1]
            if addr == UNDEF:
                arr = np.zeros(count, dtype=dt)
            else:
                arr = np.frombuffer(self._d, dtype=dt, count=count, offset=self._a(addr))
        else:
            btree, cdims = ds.layout[1], ds.layout[2]
            out = np.zeros(shape, dtype=dt)
            if btree != UNDEF:
                for offs, size, mask, addr in self._chunks(btree, len(shape)):
                    a = self._a(addr)
                    raw = self._unfilter(self._d[a:a + size], ds.filters, mask, dt.itemsize)
                    chunk = np.frombuffer(raw, dtype=dt, count=int(np.prod(cdims))).reshape(cdims)
                    sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offs, cdims, shape))
                    out[sl] = chunk[tuple(slice(0, x.stop - x.start) for x in sl)]
            arr = out
        return np.array(arr, dtype=dt.newbyteorder("=")).reshape(shape)
