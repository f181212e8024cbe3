"""Minimal HDF5 writer for test fixtures (no h5py here): superblock 0,
version-1 object headers, symbol-table groups (one leaf B-tree node, one
symbol-table node, one local heap per group), fixed-point and IEEE
floating-point datasets in either byte order, contiguous or chunked (one
leaf chunk B-tree) with the shuffle and deflate filters.  Written from the
HDF5 File Format Specification; the library's readers accept files of this
shape.  Used only to produce inputs for ``feelmri_amd.hdf5_lite`` tests.
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF


def _pad8(b: bytes) -> bytes:
    return b + b"\0" * (-len(b) % 8)


class H5Writer:
    def __init__(self):
        self.buf = bytearray(b"\0" * 96)          # superblock + root entry, filled at close
        self.tree = {}                              # nested dict: name -> dict | spec

    def create_dataset(self, path: str, data, chunks=None, shuffle=False, deflate=None, byteorder="<"):
        parts = [p for p in path.split("/") if p]
        node = self.tree
        for p in parts[:-1]:
            node = node.setdefault(p, {})
        a = np.asarray(data)
        a = a.astype(a.dtype.newbyteorder(byteorder))
        node[parts[-1]] = dict(data=a, chunks=chunks, shuffle=shuffle, deflate=deflate)

    # -- allocation ------------------------------------------------------------------
    def _put(self, b: bytes) -> int:
        while len(self.buf) % 8:
            self.buf.append(0)
        addr = len(self.buf)
        self.buf += b
        return addr

    def _patch(self, addr: int, b: bytes) -> None:
        self.buf[addr:addr + len(b)] = b

    @staticmethod
    def _header(messages) -> bytes:
        body = b"".join(struct.pack("<HHB3x", t, len(_pad8(m)), 0) + _pad8(m) for t, m in messages)
        return struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(body)) + body

    @staticmethod
    def _dtype_msg(dt: np.dtype) -> bytes:
        order = 1 if dt.byteorder == ">" or (dt.byteorder == "=" and np.little_endian is False) else 0
        n = dt.itemsize
        if dt.kind in "iu":
            bits = order | (0x08 if dt.kind == "i" else 0)
            return struct.pack("<B3BI", (1 << 4) | 0, bits, 0, 0, n) + struct.pack("<HH", 0, 8 * n)
        if dt.kind == "f":
            exp = {2: (10, 5, 10, 15), 4: (23, 8, 23, 127), 8: (52, 11, 52, 1023)}[n]
            bits0 = order | 0x20                               # mantissa normalisation: implied msb
            return (struct.pack("<B3BI", (1 << 4) | 1, bits0, 8 * n - 1, 0, n)
                    + struct.pack("<HHBBBBI", 0, 8 * n, exp[0], exp[1], 0, exp[2], exp[3]))
        raise ValueError(dt)

    def _dataset(self, spec) -> int:
        a = spec["data"]
        dt = a.dtype
        msgs = [(0x0001, struct.pack("<BBB5x", 1, a.ndim, 0) + b"".join(struct.pack("<Q", s) for s in a.shape)),
                (0x0003, self._dtype_msg(dt))]
        if spec["chunks"] is None:
            addr = self._put(a.tobytes())
            msgs.append((0x0008, struct.pack("<BBQQ", 3, 1, addr, a.nbytes)))
        else:
            cdims = tuple(spec["chunks"])
            filters = []
            if spec["shuffle"]:
                filters.append((2, [dt.itemsize]))
            if spec["deflate"] is not None:
                filters.append((1, [int(spec["deflate"])]))
            entries = []
            for idx in np.ndindex(*[-(-s // c) for s, c in zip(a.shape, cdims)]):
                offs = [i * c for i, c in zip(idx, cdims)]
                chunk = np.zeros(cdims, dtype=dt)
                sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offs, cdims, a.shape))
                chunk[tuple(slice(0, x.stop - x.start) for x in sl)] = a[sl]
                raw = chunk.tobytes()
                for fid, vals in filters:
                    if fid == 2:
                        raw = np.frombuffer(raw, np.uint8).reshape(-1, dt.itemsize).T.tobytes()
                    else:
                        raw = zlib.compress(raw, vals[0])
                entries.append((offs, len(raw), self._put(raw)))
            nd = a.ndim
            node = b"TREE" + struct.pack("<BBHQQ", 1, 0, len(entries), UNDEF, UNDEF)
            for offs, size, caddr in entries:
                node += struct.pack("<II", size, 0) + b"".join(struct.pack("<Q", o) for o in (*offs, 0))
                node += struct.pack("<Q", caddr)
            node += struct.pack("<II", 0, 0) + b"".join(struct.pack("<Q", o) for o in (*a.shape, 0))
            btree = self._put(node)
            msgs.append((0x0008, struct.pack("<BBBQ", 3, 2, nd + 1, btree)
                         + b"".join(struct.pack("<I", c) for c in (*cdims, dt.itemsize))))
            if filters:
                fp = struct.pack("<BB6x", 1, len(filters))
                for fid, vals in filters:
                    fp += struct.pack("<HHHH", fid, 0, 0, len(vals)) + b"".join(struct.pack("<I", v) for v in vals)
                    if len(vals) % 2:
                        fp += b"\0" * 4
                msgs.append((0x000B, fp))
        return self._put(self._header(msgs))

    def _group(self, tree: dict):
        """(header address, B-tree address, heap address) of a group."""
        names = sorted(tree)
        children = []
        for n in names:
            v = tree[n]
            if isinstance(v, dict) and "data" not in v:
                children.append((n, *self._group(v), 1))
            else:
                children.append((n, self._dataset(v), 0, 0, 0))
        heap_data = b"\0" * 8                               # offset 0: the empty name
        offsets = []
        for n, *_ in children:
            offsets.append(len(heap_data))
            heap_data += _pad8(n.encode() + b"\0")
        seg = self._put(heap_data)
        heap = self._put(b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap_data), UNDEF, seg))
        snod = b"SNOD" + struct.pack("<BBH", 1, 0, len(children))
        for off, (n, hdr, bt, hp, cache) in zip(offsets, children):
            snod += struct.pack("<QQI4xQQ", off, hdr, cache, bt, hp)
        snod_addr = self._put(snod)
        tree_node = b"TREE" + struct.pack("<BBHQQ", 0, 0, 1, UNDEF, UNDEF)
        tree_node += struct.pack("<QQQ", 0, snod_addr, offsets[-1] if offsets else 0)
        btree = self._put(tree_node)
        hdr = self._put(self._header([(0x0011, struct.pack("<QQ", btree, heap))]))
        return hdr, btree, heap

    def save(self, path) -> None:
        hdr, btree, heap = self._group(self.tree)
        sb = b"\x89HDF\r\n\x1a\n" + struct.pack("<BBBBBBBB", 0, 0, 0, 0, 0, 8, 8, 0)
        sb += struct.pack("<HHI", 4, 16, 0)
        sb += struct.pack("<QQQQ", 0, UNDEF, len(self.buf), UNDEF)
        sb += struct.pack("<QQI4xQQ", 0, hdr, 1, btree, heap)
        self._patch(0, sb)
        with open(path, "wb") as f:
            f.write(bytes(self.buf))
