"""A small HDF5 writer and reader for particle-field files (numpy only).

FLOWVPM's `save` writes one HDF5 file per step holding flat datasets in the
root group (np, nt, t, X, Gamma, sigma, circulation, vol, static, i,
connectivity) next to an XDMF descriptor (src/FLOWUnsteady_simulation.jl:264
reads such a file back through `vpm.read!`).  No HDF5 binding is importable
here, so this module implements the part of the HDF5 file format (HDF Group,
"HDF5 File Format Specification Version 3.0") those files use:

writer  superblock version 0, root group as a symbol table (version 1 B-tree
        node + symbol table node + local heap), version 1 object headers, one
        contiguous dataset per array: dataspace (version 2, scalar or simple), datatype
        (IEEE float or two's-complement integer, little-endian), fill value
        (version 2, undefined), data layout (version 3, contiguous).
reader  superblocks 0-3; object headers version 1 and 2 with continuation
        blocks; groups as symbol tables or link messages; dataspaces 1-2;
        fixed-point and floating-point datatypes of either byte order;
        compact, contiguous and chunked layouts (version 3 with a version 1
        B-tree; version 4 contiguous, compact and single-chunk); no filters.

Only flat groups of datasets are supported: that is the content of a
particle-field file.
"""
from __future__ import annotations

import struct

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF
GROUP_LEAF_K = 16        # symbol table node capacity 2K written into the superblock
GROUP_INTERNAL_K = 16


def _pad8(n: int) -> int:
    return (n + 7) & ~7


# ---------------------------------------------------------------------------
# writer
# ---------------------------------------------------------------------------
def _datatype_message(dt: np.dtype) -> bytes:
    dt = np.dtype(dt).newbyteorder("<")
    if dt.kind == "f" and dt.itemsize in (4, 8):
        # class 1, version 1; byte order LE, implied-MSB mantissa, sign bit
        bits = dt.itemsize * 8
        if bits == 64:
            eloc, esize, msize, bias = 52, 11, 52, 1023
        else:
            eloc, esize, msize, bias = 23, 8, 23, 127
        head = bytes([0x11, 0x20, bits - 1, 0x00]) + struct.pack("<I", dt.itemsize)
        props = struct.pack("<HHBBBBI", 0, bits, eloc, esize, 0, msize, bias)
        return head + props
    if dt.kind in "iu" and dt.itemsize in (1, 2, 4, 8):
        signed = 0x08 if dt.kind == "i" else 0x00
        head = bytes([0x10, signed, 0x00, 0x00]) + struct.pack("<I", dt.itemsize)
        return head + struct.pack("<HH", 0, dt.itemsize * 8)
    raise TypeError(f"unsupported dtype {dt}")


def _dataspace_message(shape) -> bytes:
    # version 2 (dimensionality, flags, type: 0 scalar / 1 simple), no maximum
    # dimensions
    out = bytes([2, len(shape), 0, 0 if len(shape) == 0 else 1])
    for d in shape:
        out += struct.pack("<Q", int(d))
    return out


def _messages_v1(msgs) -> bytes:
    """Version 1 object header: 16-byte prefix, messages padded to 8 bytes."""
    body = b""
    for mtype, data in msgs:
        size = _pad8(len(data))
        body += struct.pack("<HHB3x", mtype, size, 0) + data + b"\x00" * (size - len(data))
    prefix = struct.pack("<BBHII", 1, 0, len(msgs), 1, len(body)) + b"\x00" * 4
    return prefix + body


def write(path: str, datasets: dict) -> None:
    """Write `datasets` ({name: numpy array or scalar}) as the datasets of the
    root group of a new HDF5 file at `path`.  Arrays are stored C-ordered
    (row-major) with their numpy shape as the HDF5 dimensions."""
    names = sorted(datasets)                    # symbol table nodes keep names sorted
    if len(names) > 2 * GROUP_LEAF_K:
        raise ValueError(f"at most {2 * GROUP_LEAF_K} datasets per file")
    arrays = {}
    for k in names:
        a = np.asarray(datasets[k])
        if a.dtype.kind == "b":
            a = a.astype(np.uint8)
        # (ascontiguousarray would turn a 0-d array into a 1-d one)
        arrays[k] = np.array(a, dtype=a.dtype.newbyteorder("<"), order="C", copy=True)

    # local heap data segment: "" at offset 0, then the names
    heap = bytearray(b"\x00" * 8)
    name_off = {}
    for k in names:
        name_off[k] = len(heap)
        enc = k.encode("ascii") + b"\x00"
        heap += enc + b"\x00" * (_pad8(len(enc)) - len(enc))

    SB, ROOT_OH = 0, 96
    root_oh = _messages_v1([(0x0011, b"\x00" * 16)])          # patched below
    heap_hdr = ROOT_OH + len(root_oh)
    heap_data = heap_hdr + 32
    btree = heap_data + len(heap)
    btree_size = 24 + (2 * GROUP_INTERNAL_K + 1) * 8 + 2 * GROUP_INTERNAL_K * 8
    snod = btree + btree_size
    snod_size = 8 + 2 * GROUP_LEAF_K * 40
    pos = snod + snod_size

    # dataset object headers, then their data
    headers, hdr_addr = {}, {}
    for k in names:
        a = arrays[k]
        msgs = [(0x0001, _dataspace_message(a.shape)), (0x0003, _datatype_message(a.dtype)),
                (0x0005, bytes([2, 2, 2, 0])),
                (0x0008, b"\x00" * 18)]                      # layout, patched below
        headers[k] = msgs
        hdr_addr[k] = pos
        pos += len(_messages_v1(msgs))
    data_addr = {}
    for k in names:
        pos = _pad8(pos)
        data_addr[k] = pos
        pos += arrays[k].nbytes
    eof = _pad8(pos)

    buf = bytearray(eof)
    # superblock version 0
    sb = SIGNATURE + bytes([0, 0, 0, 0, 0, 8, 8, 0])
    sb += struct.pack("<HHI", GROUP_LEAF_K, GROUP_INTERNAL_K, 0)
    sb += struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF)
    sb += struct.pack("<QQII", 0, ROOT_OH, 1, 0) + struct.pack("<QQ", btree, heap_hdr)
    assert len(sb) == 96
    buf[SB:SB + 96] = sb
    root_oh = _messages_v1([(0x0011, struct.pack("<QQ", btree, heap_hdr))])
    buf[ROOT_OH:ROOT_OH + len(root_oh)] = root_oh
    # local heap: no free block (offset 1 is the null free-list offset)
    buf[heap_hdr:heap_hdr + 32] = b"HEAP" + bytes([0, 0, 0, 0]) + struct.pack(
        "<QQQ", len(heap), 1, heap_data)
    buf[heap_data:heap_data + len(heap)] = heap
    # group B-tree: one leaf node with one child (the symbol table node)
    bt = b"TREE" + bytes([0, 0]) + struct.pack("<H", 1) + struct.pack("<QQ", UNDEF, UNDEF)
    bt += struct.pack("<QQQ", 0, snod, name_off[names[-1]] if names else 0)
    buf[btree:btree + len(bt)] = bt
    # symbol table node
    sn = b"SNOD" + bytes([1, 0]) + struct.pack("<H", len(names))
    for k in names:
        sn += struct.pack("<QQII", name_off[k], hdr_addr[k], 0, 0) + b"\x00" * 16
    buf[snod:snod + len(sn)] = sn
    # dataset headers and data
    for k in names:
        a = arrays[k]
        msgs = headers[k][:3] + [(0x0008, bytes([3, 1]) + struct.pack("<QQ", data_addr[k],
                                                                     a.nbytes))]
        oh = _messages_v1(msgs)
        buf[hdr_addr[k]:hdr_addr[k] + len(oh)] = oh
        buf[data_addr[k]:data_addr[k] + a.nbytes] = a.tobytes()
    with open(path, "wb") as f:
        f.write(bytes(buf))


# ---------------------------------------------------------------------------
# reader
# ---------------------------------------------------------------------------
class FormatError(ValueError):
    pass


class _File:
    def __init__(self, data: bytes):
        self.d = data
        base = self._find_superblock()
        self.base = base
        v = data[base + 8]
        if v in (0, 1):
            self.so, self.sl = data[base + 13], data[base + 14]
            if (self.so, self.sl) != (8, 8):
                raise FormatError("only 8-byte offsets and lengths are supported")
            p = base + 24 + (4 if v == 1 else 0)
            self.base_addr = self.u(p, 8)
            # root symbol table entry after the four addresses
            ent = p + 32
            self.root = ("oh", self.u(ent + 8, 8))
        elif v in (2, 3):
            self.so, self.sl = data[base + 9], data[base + 10]
            if (self.so, self.sl) != (8, 8):
                raise FormatError("only 8-byte offsets and lengths are supported")
            self.base_addr = self.u(base + 12, 8)
            self.root = ("oh", self.u(base + 12 + 24, 8))
        else:
            raise FormatError(f"superblock version {v}")

    def _find_superblock(self):
        off = 0
        while off + 8 <= len(self.d):
            if self.d[off:off + 8] == SIGNATURE:
                return off
            off = 512 if off == 0 else off * 2
        raise FormatError("no HDF5 signature")

    def u(self, p, n):
        return int.from_bytes(self.d[p:p + n], "little")

    def addr(self, a):
        return self.base_addr + a

    # -- object headers ------------------------------------------------------------
    def messages(self, oh):
        p = self.addr(oh)
        out = []
        if self.d[p:p + 4] == b"OHDR":
            ver, flags = self.d[p + 4], self.d[p + 5]
            q = p + 6
            if flags & 0x20:
                q += 16
            if flags & 0x10:
                q += 4
            sz_len = 1 << (flags & 3)
            size = self.u(q, sz_len)
            q += sz_len
            # chunk #0: messages (and a gap shorter than a message header)
            # in [q, q + size), then its checksum; continuation blocks:
            # "OCHK", messages, checksum
            blocks = [(q, q + size)]
            hdr = 4 + (2 if flags & 0x04 else 0)
            while blocks:
                s, e = blocks.pop(0)
                while s + hdr <= e:
                    mtype, msize = self.d[s], self.u(s + 1, 2)
                    s += hdr
                    out.append((mtype, s, msize))
                    if mtype == 0x10:
                        ca, cl = self.u(s, 8), self.u(s + 8, 8)
                        cp = self.addr(ca)
                        if self.d[cp:cp + 4] != b"OCHK":
                            raise FormatError("bad continuation block")
                        blocks.append((cp + 4, cp + cl - 4))
                    s += msize
            return out
        if self.d[p] != 1:
            raise FormatError(f"object header version {self.d[p]}")
        nmsg, size = self.u(p + 2, 2), self.u(p + 8, 4)
        blocks = [(p + 16, p + 16 + size)]
        while blocks and len(out) < nmsg:
            s, e = blocks.pop(0)
            while s + 8 <= e and len(out) < nmsg:
                mtype, msize = self.u(s, 2), self.u(s + 2, 2)
                out.append((mtype, s + 8, msize))
                if mtype == 0x10:
                    ca, cl = self.u(s + 8, 8), self.u(s + 16, 8)
                    blocks.append((self.addr(ca), self.addr(ca) + cl))
                s += 8 + msize
        return out

    # -- groups ----------------------------------------------------------------------
    def members(self, oh):
        """{name: object header address} of the group whose header is at oh."""
        out = {}
        for mtype, p, _ in self.messages(oh):
            if mtype == 0x11:                   # symbol table
                self._btree_group(self.u(p, 8), self.u(p + 8, 8), out)
            elif mtype == 0x06:                 # link (compact storage)
                name, target = self._link(p)
                if target is not None:
                    out[name] = target
            elif mtype == 0x02:                 # link info: dense storage
                q = p + 2 + (8 if self.d[p + 1] & 1 else 0)
                fheap, bt = self.u(q, 8), self.u(q + 8, 8)
                if fheap != UNDEF:
                    self._dense_links(fheap, bt, out)
        return out

    # dense link storage: link messages in a fractal heap, found through the
    # heap IDs of the name-index version 2 B-tree (record type 5: name hash,
    # heap ID)
    def _dense_links(self, fheap, bt, out):
        p = self.addr(bt)
        if self.d[p:p + 4] != b"BTHD":
            raise FormatError("bad v2 B-tree header")
        btype, rsize, depth = self.d[p + 5], self.u(p + 10, 2), self.u(p + 12, 2)
        root, nroot = self.u(p + 16, 8), self.u(p + 24, 2)
        if btype != 5:
            raise FormatError(f"v2 B-tree type {btype} for links")
        if depth != 0:
            raise FormatError("link name index deeper than one leaf is not supported")
        heap = self._fheap(fheap)
        q = self.addr(root)
        if self.d[q:q + 4] != b"BTLF":
            raise FormatError("bad v2 B-tree leaf")
        q += 6
        for i in range(nroot):
            rec = q + i * rsize
            hid = self.d[rec + 4:rec + rsize]
            name, target = self._link(heap(hid))
            if target is not None:
                out[name] = target

    def _fheap(self, fheap):
        """A function mapping a managed-object heap ID to the object's file
        position in the fractal heap whose header is at fheap."""
        p = self.addr(fheap)
        if self.d[p:p + 4] != b"FRHP":
            raise FormatError("bad fractal heap header")
        q = p + 5
        q += 2 + 2 + 1            # heap ID length, filter length, flags
        max_obj = self.u(q, 4)
        q += 4 + 8 + 8            # max managed size, next huge id, huge B-tree
        q += 8 + 8 + 8 + 8 + 8 + 8 + 8 + 8 + 8 + 8
        width, start = self.u(q, 2), self.u(q + 2, 8)
        max_direct, max_bits = self.u(q + 10, 8), self.u(q + 18, 2)
        root, rows = self.u(q + 22, 8), self.u(q + 30, 2)
        off_size = (max_bits + 7) // 8
        len_size = (min(max_direct, max_obj).bit_length() + 7) // 8
        # direct blocks: (heap offset, file address) of every block
        blocks = []
        if rows == 0:
            blocks.append((0, root))
        else:
            b = self.addr(root)
            if self.d[b:b + 4] != b"FHIB":
                raise FormatError("bad fractal heap indirect block")
            e = b + 5 + 8 + off_size
            hoff = 0
            for r in range(rows):
                bsize = start * (1 << max(0, r - 1))
                if bsize > max_direct:
                    raise FormatError("nested indirect blocks are not supported")
                for _ in range(width):
                    a = self.u(e, 8)
                    e += 8
                    if a != UNDEF:
                        blocks.append((hoff, a))
                    hoff += bsize

        def locate(hid):
            if (hid[0] >> 4) & 3 != 0:
                raise FormatError("huge or tiny heap objects are not supported")
            off = int.from_bytes(hid[1:1 + off_size], "little")
            ln = int.from_bytes(hid[1 + off_size:1 + off_size + len_size], "little")
            base = max((bo for bo in blocks if bo[0] <= off), key=lambda bo: bo[0])
            return self.addr(base[1]) + (off - base[0])

        return locate

    def _link(self, p):
        ver, flags = self.d[p], self.d[p + 1]
        if ver != 1:
            raise FormatError(f"link message version {ver}")
        q = p + 2
        ltype = 0
        if flags & 0x08:
            ltype = self.d[q]
            q += 1
        if flags & 0x04:
            q += 8
        if flags & 0x10:
            q += 1
        nlen_size = 1 << (flags & 3)
        nlen = self.u(q, nlen_size)
        q += nlen_size
        name = self.d[q:q + nlen].decode("utf-8")
        q += nlen
        return name, (self.u(q, 8) if ltype == 0 else None)

    def _heap_name(self, heap, off):
        p = self.addr(heap)
        if self.d[p:p + 4] != b"HEAP":
            raise FormatError("bad local heap")
        data = self.addr(self.u(p + 24, 8))
        end = self.d.index(b"\x00", data + off)
        return self.d[data + off:end].decode("utf-8")

    def _btree_group(self, bt, heap, out):
        p = self.addr(bt)
        if self.d[p:p + 4] != b"TREE" or self.d[p + 4] != 0:
            raise FormatError("bad group B-tree node")
        level, n = self.d[p + 5], self.u(p + 6, 2)
        q = p + 24 + 8                      # key 0, then (child, key) pairs
        for i in range(n):
            child = self.u(q + 16 * i, 8)
            if level > 0:
                self._btree_group(child, heap, out)
                continue
            s = self.addr(child)
            if self.d[s:s + 4] != b"SNOD":
                raise FormatError("bad symbol table node")
            for j in range(self.u(s + 6, 2)):
                e = s + 8 + 40 * j
                out[self._heap_name(heap, self.u(e, 8))] = self.u(e + 8, 8)

    # -- datasets --------------------------------------------------------------------
    def dataset(self, oh):
        shape = dtype = layout = None
        for mtype, p, size in self.messages(oh):
            if mtype == 0x01:
                shape = self._dataspace(p)
            elif mtype == 0x03:
                dtype = self._datatype(p)
            elif mtype == 0x08:
                layout = (p, size)
            elif mtype == 0x0B:
                raise FormatError("filtered (compressed) datasets are not supported")
        if shape is None or dtype is None or layout is None:
            raise FormatError("not a dataset")
        count = int(np.prod(shape)) if shape else 1
        raw = self._layout_bytes(layout[0], shape, dtype.itemsize, count)
        a = np.frombuffer(raw, dtype=dtype, count=count).copy()
        return a.reshape(shape) if shape else a.reshape(())

    def _dataspace(self, p):
        ver, rank, flags = self.d[p], self.d[p + 1], self.d[p + 2]
        if ver == 1:
            q = p + 8
        elif ver == 2:
            if self.d[p + 3] == 2:          # null dataspace
                return (0,)
            q = p + 4
        else:
            raise FormatError(f"dataspace version {ver}")
        return tuple(self.u(q + 8 * i, 8) for i in range(rank))

    def _datatype(self, p):
        cls, bits0 = self.d[p] & 0x0F, self.d[p + 1]
        size = self.u(p + 4, 4)
        order = ">" if bits0 & 1 else "<"
        if cls == 0:
            return np.dtype(f"{order}{'i' if bits0 & 0x08 else 'u'}{size}")
        if cls == 1:
            return np.dtype(f"{order}f{size}")
        raise FormatError(f"datatype class {cls}")

    def _layout_bytes(self, p, shape, esize, count):
        ver, cls = self.d[p], self.d[p + 1]
        nbytes = count * esize
        if ver == 3:
            if cls == 0:                    # compact
                n = self.u(p + 2, 2)
                return self.d[p + 4:p + 4 + n]
            if cls == 1:                    # contiguous
                a = self.u(p + 2, 8)
                if a == UNDEF:
                    return b"\x00" * nbytes
                return self.d[self.addr(a):self.addr(a) + nbytes]
            if cls == 2:                    # chunked, version 1 B-tree
                nd = self.d[p + 2]
                bt = self.u(p + 3, 8)
                cdims = [self.u(p + 11 + 4 * i, 4) for i in range(nd - 1)]
                return self._chunked(bt, nd, shape, cdims, esize)
        elif ver == 4:
            if cls == 0:
                n = self.u(p + 2, 2)
                return self.d[p + 4:p + 4 + n]
            if cls == 1:
                a = self.u(p + 2, 8)
                return self.d[self.addr(a):self.addr(a) + nbytes]
            if cls == 2:
                flags, nd, w = self.d[p + 2], self.d[p + 3], self.d[p + 4]
                cdims = [self.u(p + 5 + w * i, w) for i in range(nd)]
                q = p + 5 + w * nd
                itype = self.d[q]
                if itype == 1:              # single chunk
                    q += 1
                    if flags & 2:
                        raise FormatError("filtered single chunk")
                    a = self.u(q, 8)
                    return self.d[self.addr(a):self.addr(a) + nbytes]
                raise FormatError(f"chunk index type {itype} is not supported")
        raise FormatError(f"layout version {ver} class {cls}")

    def _chunked(self, bt, nd, shape, cdims, esize):
        out = np.zeros(int(np.prod(shape)) * esize, dtype=np.uint8).reshape(
            tuple(shape) + (esize,))
        cshape = tuple(cdims)

        def walk(a):
            p = self.addr(a)
            if self.d[p:p + 4] != b"TREE" or self.d[p + 4] != 1:
                raise FormatError("bad chunk B-tree node")
            level, n = self.d[p + 5], self.u(p + 6, 2)
            ksize = 8 + 8 * nd
            q = p + 24
            for i in range(n):
                k = q + i * (ksize + 8)
                csize, fmask = self.u(k, 4), self.u(k + 4, 4)
                offs = [self.u(k + 8 + 8 * j, 8) for j in range(nd - 1)]
                child = self.u(k + ksize, 8)
                if level > 0:
                    walk(child)
                    continue
                if fmask:
                    raise FormatError("filtered chunk")
                c = np.frombuffer(self.d, dtype=np.uint8, count=csize,
                                  offset=self.addr(child)).reshape(cshape + (esize,))
                sl = tuple(slice(o, min(o + cd, s)) for o, cd, s in zip(offs, cdims, shape))
                sub = tuple(slice(0, x.stop - x.start) for x in sl)
                out[sl] = c[sub]

        walk(bt)
        return out.tobytes()


def read(path: str, names=None) -> dict:
    """{name: numpy array} of the datasets in the root group of the HDF5 file
    at `path` (all of them, or those in `names`).  Scalars come back as 0-d
    arrays."""
    with open(path, "rb") as f:
        fl = _File(f.read())
    kind, oh = fl.root
    members = fl.members(oh)
    out = {}
    for name, addr in members.items():
        if names is not None and name not in names:
            continue
        try:
            out[name] = fl.dataset(addr)
        except FormatError:
            if names is not None:
                raise
    return out
