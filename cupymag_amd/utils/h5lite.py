"""A small HDF5 writer and reader for the magnetization checkpoint.

``write_array`` in the reference (magnetization_io.py:21-41) stores the
array with h5py as dataset ``data`` (gzip level 6) and the root attributes
``shape`` and ``dtype``.  h5py is not part of this image, so this module
writes that file itself, following the HDF5 file format specification
(version 0 superblock, version 1 object headers, a symbol-table root group,
a chunked dataset indexed by a version 1 B-tree, the deflate filter), and
reads back the subset of the format such files use: superblock 0/1,
version 1 object headers with continuation blocks, symbol-table groups,
contiguous / compact / B-tree-chunked layouts with deflate and shuffle
filters, little- and big-endian integer and IEEE float data.

Differences from an h5py-written file that do not change what a reader
gets for ``data``: the ``dtype`` attribute is a fixed-length ASCII string
(h5py writes a variable-length one), and chunk shapes are this module's
own choice (whole rows, at most 64 chunks).
"""
import struct
import zlib

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF
_GROUP_LEAF_K = 4          # symbol table node holds 2K entries
_GROUP_INTERNAL_K = 16     # group B-tree node holds 2K children
_CHUNK_K = 32              # indexed-storage B-tree node holds 2K children (v0 superblock default)
_HEAP_FREE_NULL = 1        # "no free block" in a local heap's free list


def _pad8(b):
    return b + b"\0" * (-len(b) % 8)


# ---------------------------------------------------------------- messages

def _dt_float64():
    # class 1 (floating point), version 1; LE, implied-msb mantissa, sign bit 63
    return struct.pack("<B3sI", 0x11, bytes([0x20, 63, 0]), 8) + struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)


def _dt_int64():
    # class 0 (fixed point), version 1; LE, signed
    return struct.pack("<B3sI", 0x10, bytes([0x08, 0, 0]), 8) + struct.pack("<HH", 0, 64)


def _dt_string(n):
    # class 3 (string), version 1; null-padded ASCII
    return struct.pack("<B3sI", 0x13, bytes([0x01, 0, 0]), n)


def _dataspace(dims, maxdims=False):
    flags = 1 if maxdims else 0
    b = struct.pack("<BBBB4x", 1, len(dims), flags, 0)
    b += b"".join(struct.pack("<Q", d) for d in dims)
    if maxdims:
        b += b"".join(struct.pack("<Q", d) for d in dims)
    return b


def _attribute(name, dtype_msg, space_msg, data):
    nm = name.encode() + b"\0"
    return (struct.pack("<BBHHH", 1, 0, len(nm), len(dtype_msg), len(space_msg)) + _pad8(nm) + _pad8(dtype_msg)
            + _pad8(space_msg) + data)


def _object_header(messages):
    """Version 1 object header: prefix (aligned to 8) + messages."""
    body = b""
    for mtype, data in messages:
        data = _pad8(data)
        body += struct.pack("<HHB3x", mtype, len(data), 0) + data
    return struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(body)) + body


# ---------------------------------------------------------------- writer

def write(filename, data, attrs=None, level=6):
    """Write ``data`` (a numeric array of rank >= 1) as dataset ``data``
    with deflate ``level`` and root attributes ``attrs`` (name -> tuple of
    ints, or str)."""
    data = np.ascontiguousarray(data)
    if data.dtype != np.float64:
        data = data.astype(np.float64)
    if data.ndim == 0:
        data = data.reshape(1)
    dims = list(data.shape)
    rows = dims[0]
    row_elems = int(np.prod(dims[1:])) if len(dims) > 1 else 1
    nchunk = max(1, min(64, -(-rows // max(1, (1 << 20) // max(1, row_elems)))))
    crows = max(1, -(-rows // nchunk))
    nchunk = max(1, -(-rows // crows))
    chunk_dims = [crows] + dims[1:]
    flat = data.reshape(rows, row_elems)

    out = bytearray(b"\0" * 96)  # superblock, written last

    def alloc(b):
        addr = len(out)
        out.extend(b)
        out.extend(b"\0" * (-len(out) % 8))
        return addr

    # chunks (each a full chunk: the edge chunk is zero-padded)
    chunk_recs = []
    for c in range(nchunk):
        blk = np.zeros((crows, row_elems), dtype="<f8")
        part = flat[c * crows:(c + 1) * crows]
        blk[:len(part)] = part
        comp = zlib.compress(blk.tobytes(), level)
        chunk_recs.append((c * crows, len(comp), alloc(comp)))

    # chunk B-tree (one leaf node, allocated at full size)
    rank1 = len(dims) + 1

    def ckey(nbytes, offs):
        return struct.pack("<II", nbytes, 0) + b"".join(struct.pack("<Q", o) for o in offs)

    node = b"TREE" + struct.pack("<BBHQQ", 1, 0, nchunk, UNDEF, UNDEF)
    for (off, nbytes, addr) in chunk_recs:
        node += ckey(nbytes, [off] + [0] * (len(dims) - 1) + [0]) + struct.pack("<Q", addr)
    last = chunk_recs[-1][0] + crows
    node += ckey(0, [last] + dims[1:] + [0])
    keysz = 8 + 8 * rank1
    full = 24 + (2 * _CHUNK_K + 1) * keysz + 2 * _CHUNK_K * 8
    btree_addr = alloc(node + b"\0" * (full - len(node)))

    # dataset object header
    pipeline = struct.pack("<BB6x", 1, 1) + struct.pack("<HHHH", 1, 0, 0, 1) + struct.pack("<I4x", level)
    layout = struct.pack("<BBBQ", 3, 2, rank1, btree_addr) + b"".join(struct.pack("<I", d) for d in chunk_dims) \
        + struct.pack("<I", 8)
    fill = struct.pack("<BBBB", 2, 3, 2, 0)
    dset_addr = alloc(_object_header([(0x0001, _dataspace(dims, maxdims=True)), (0x0003, _dt_float64()),
                                      (0x0005, fill), (0x0008, layout), (0x000B, pipeline)]))

    # root group: local heap with the link name, symbol table node, group B-tree
    names = _pad8(b"\0") + _pad8(b"data\0")
    heap_size = len(names) + 16 + 48
    free = struct.pack("<QQ", _HEAP_FREE_NULL, heap_size - len(names))
    heap_data = names + free + b"\0" * (heap_size - len(names) - len(free))
    heap_data_addr = alloc(heap_data)
    heap_addr = alloc(b"HEAP" + struct.pack("<B3xQQQ", 0, heap_size, len(names), heap_data_addr))

    entry = struct.pack("<QQI4x16x", 8, dset_addr, 0)
    snod = b"SNOD" + struct.pack("<BBH", 1, 0, 1) + entry
    snod += b"\0" * (8 + 2 * _GROUP_LEAF_K * 40 - len(snod))
    snod_addr = alloc(snod)
    gnode = b"TREE" + struct.pack("<BBHQQ", 0, 0, 1, UNDEF, UNDEF) + struct.pack("<QQQ", 0, snod_addr, 8)
    gfull = 24 + (2 * _GROUP_INTERNAL_K + 1) * 8 + 2 * _GROUP_INTERNAL_K * 8
    gtree_addr = alloc(gnode + b"\0" * (gfull - len(gnode)))

    msgs = [(0x0011, struct.pack("<QQ", gtree_addr, heap_addr))]
    for name, val in (attrs or {}).items():
        if isinstance(val, str):
            raw = val.encode()
            msgs.append((0x000C, _attribute(name, _dt_string(len(raw)), _dataspace([]), raw)))
        else:
            arr = np.asarray(val, dtype="<i8").reshape(-1)
            msgs.append((0x000C, _attribute(name, _dt_int64(), _dataspace([len(arr)]), arr.tobytes())))
    root_addr = alloc(_object_header(msgs))

    eof = len(out)
    sb = SIGNATURE + struct.pack("<BBBBBBBB", 0, 0, 0, 0, 0, 8, 8, 0)
    sb += struct.pack("<HHI", _GROUP_LEAF_K, _GROUP_INTERNAL_K, 0)
    sb += struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF)
    sb += struct.pack("<QQII", 0, root_addr, 1, 0) + struct.pack("<QQ", gtree_addr, heap_addr)
    assert len(sb) == 96
    out[0:96] = sb
    with open(filename, "wb") as f:
        f.write(bytes(out))


# ---------------------------------------------------------------- reader

class H5FormatError(ValueError):
    pass


class _File:
    def __init__(self, buf):
        self.b = buf
        if buf[:8] != SIGNATURE:
            raise H5FormatError("not an HDF5 file (signature)")
        ver = buf[8]
        if ver not in (0, 1):
            raise H5FormatError(f"superblock version {ver} is not supported")
        self.so, self.sl = buf[13], buf[14]
        if self.so != 8 or self.sl != 8:
            raise H5FormatError("only 8-byte offsets and lengths are supported")
        self.chunk_k = _CHUNK_K
        p = 24
        if ver == 1:
            self.chunk_k = struct.unpack_from("<H", buf, 24)[0]
            p = 28
        self.base = struct.unpack_from("<Q", buf, p)[0]
        self.root = struct.unpack_from("<Q", buf, p + 32 + 8)[0]

    def u(self, fmt, off):
        return struct.unpack_from(fmt, self.b, off)

    def messages(self, addr):
        """(type, data bytes) of a version 1 object header, following
        continuation messages."""
        b = self.b
        if b[addr] != 1:
            raise H5FormatError(f"object header version {b[addr]} is not supported")
        nmsg, _, size = self.u("<HII", addr + 2)
        blocks = [(addr + 16, size)]
        out = []
        while blocks and len(out) < nmsg:
            start, length = blocks.pop(0)
            p = start
            while p + 8 <= start + length and len(out) < nmsg:
                mtype, msize, _flags = self.u("<HHB", p)
                data = bytes(b[p + 8:p + 8 + msize])
                if mtype == 0x0010:  # continuation
                    off, ln = struct.unpack_from("<QQ", data)
                    blocks.append((off, ln))
                out.append((mtype, data))
                p += 8 + msize
        return out

    def heap_string(self, heap_addr, off):
        if self.b[heap_addr:heap_addr + 4] != b"HEAP":
            raise H5FormatError("bad local heap")
        data_addr = self.u("<Q", heap_addr + 24)[0]
        s = data_addr + off
        e = self.b.index(b"\0", s)
        return self.b[s:e].decode()

    def group_entries(self, btree_addr, heap_addr):
        b = self.b
        if b[btree_addr:btree_addr + 4] != b"TREE":
            raise H5FormatError("bad group B-tree")
        ntype, level, used = self.u("<BBH", btree_addr + 4)
        p = btree_addr + 24
        res = {}
        for i in range(used):
            child = self.u("<Q", p + 8)[0]
            p += 16
            if level > 0:
                res.update(self.group_entries(child, heap_addr))
                continue
            if b[child:child + 4] != b"SNOD":
                raise H5FormatError("bad symbol table node")
            nsym = self.u("<H", child + 6)[0]
            for k in range(nsym):
                e = child + 8 + 40 * k
                name_off, oh = self.u("<QQ", e)
                res[self.heap_string(heap_addr, name_off)] = oh
        return res

    def lookup(self, name):
        for mtype, data in self.messages(self.root):
            if mtype == 0x0011:
                bt, hp = struct.unpack_from("<QQ", data)
                ents = self.group_entries(bt, hp)
                if name not in ents:
                    raise KeyError(name)
                return ents[name]
        raise H5FormatError("root group has no symbol table (newer group formats are not supported)")


def _parse_dtype(d):
    cls, ver = d[0] & 0x0F, d[0] >> 4
    bits0 = d[1]
    size = struct.unpack_from("<I", d, 4)[0]
    order = ">" if bits0 & 1 else "<"
    if cls == 1:
        return np.dtype(f"{order}f{size}")
    if cls == 0:
        return np.dtype(f"{order}{'i' if bits0 & 0x08 else 'u'}{size}")
    raise H5FormatError(f"datatype class {cls} (version {ver}) is not supported")


def _parse_space(d):
    ver, rank, flags = d[0], d[1], d[2]
    p = 8 if ver == 1 else 4
    return [struct.unpack_from("<Q", d, p + 8 * i)[0] for i in range(rank)]


def _filters(d):
    ver, n = d[0], d[1]
    p = 8 if ver == 1 else 2
    out = []
    for _ in range(n):
        fid = struct.unpack_from("<H", d, p)[0]
        if ver == 1 or fid >= 256:
            nlen, flags, ncd = struct.unpack_from("<HHH", d, p + 2)
            p += 8 + nlen + (-nlen % 8 if ver == 1 else 0)
        else:
            flags, ncd = struct.unpack_from("<HH", d, p + 2)
            nlen = 0
            p += 6
        cd = list(struct.unpack_from(f"<{ncd}I", d, p))
        p += 4 * ncd + (4 if (ver == 1 and ncd % 2) else 0)
        out.append((fid, cd))
    return out


def _unfilter(raw, filters, mask, itemsize):
    for k, (fid, cd) in reversed(list(enumerate(filters))):
        if mask & (1 << k):
            continue
        if fid == 1:
            raw = zlib.decompress(raw)
        elif fid == 2:
            n = len(raw) // itemsize
            raw = np.frombuffer(raw, dtype=np.uint8).reshape(itemsize, n).T.tobytes()
        else:
            raise H5FormatError(f"filter {fid} is not supported")
    return raw


def read(filename, name="data"):
    """Read dataset ``name`` of an HDF5 file into a numpy array."""
    with open(filename, "rb") as f:
        buf = f.read()
    F = _File(buf)
    msgs = dict()
    for mtype, data in F.messages(F.lookup(name)):
        msgs.setdefault(mtype, data)
    dims = _parse_space(msgs[0x0001])
    dt = _parse_dtype(msgs[0x0003])
    filters = _filters(msgs[0x000B]) if 0x000B in msgs else []
    lay = msgs[0x0008]
    if lay[0] != 3:
        raise H5FormatError(f"layout message version {lay[0]} is not supported")
    n = int(np.prod(dims)) if dims else 1
    cls = lay[1]
    if cls == 0:  # compact
        size = struct.unpack_from("<H", lay, 2)[0]
        return np.frombuffer(lay[4:4 + size], dtype=dt, count=n).reshape(dims).astype(dt.newbyteorder("="))
    if cls == 1:  # contiguous
        addr, size = struct.unpack_from("<QQ", lay, 2)
        if addr == UNDEF:
            return np.zeros(dims, dtype=dt.newbyteorder("="))
        return np.frombuffer(buf, dtype=dt, count=n, offset=addr).reshape(dims).astype(dt.newbyteorder("="))
    if cls != 2:
        raise H5FormatError(f"layout class {cls} is not supported")
    rank1 = lay[2]
    btree = struct.unpack_from("<Q", lay, 3)[0]
    cdims = list(struct.unpack_from(f"<{rank1}I", lay, 11))[:-1]
    out = np.zeros(dims, dtype=dt.newbyteorder("="))

    def walk(addr):
        if buf[addr:addr + 4] != b"TREE":
            raise H5FormatError("bad chunk B-tree")
        ntype, level, used = struct.unpack_from("<BBH", buf, addr + 4)
        keysz = 8 + 8 * rank1
        p = addr + 24
        for i in range(used):
            nbytes, mask = struct.unpack_from("<II", buf, p)
            offs = struct.unpack_from(f"<{rank1}Q", buf, p + 8)[:-1]
            child = struct.unpack_from("<Q", buf, p + keysz)[0]
            p += keysz + 8
            if level > 0:
                walk(child)
                continue
            raw = _unfilter(buf[child:child + nbytes], filters, mask, dt.itemsize)
            blk = np.frombuffer(raw, dtype=dt, count=int(np.prod(cdims))).reshape(cdims)
            sl = tuple(slice(o, min(o + c, d)) for o, c, d in zip(offs, cdims, dims))
            out[sl] = blk[tuple(slice(0, s.stop - s.start) for s in sl)]

    walk(btree)
    return out


def read_attrs(filename):
    """Root-group attributes (integer arrays and fixed-length strings)."""
    with open(filename, "rb") as f:
        buf = f.read()
    F = _File(buf)
    res = {}
    for mtype, d in F.messages(F.root):
        if mtype != 0x000C or d[0] != 1:
            continue
        nlen, dlen, slen = struct.unpack_from("<HHH", d, 2)
        p = 8
        name = d[p:p + nlen - 1].decode()
        p += nlen + (-nlen % 8)
        dtm = d[p:p + dlen]
        p += dlen + (-dlen % 8)
        spm = d[p:p + slen]
        p += slen + (-slen % 8)
        dims = _parse_space(spm)
        if dtm[0] & 0x0F == 3:
            size = struct.unpack_from("<I", dtm, 4)[0]
            res[name] = d[p:p + size].rstrip(b"\0").decode()
        else:
            dt = _parse_dtype(dtm)
            cnt = int(np.prod(dims)) if dims else 1
            res[name] = tuple(int(v) for v in np.frombuffer(d, dtype=dt, count=cnt, offset=p))
    return res
