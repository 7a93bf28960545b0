"""Writes UVF files holding an ExtendedOctree (test fixtures for the reader).

The layout follows the reference's writers: GlobalHeader::CopyHeaderToFile
(IO/UVF/GlobalHeader.cpp:50-57) after the "UVF-DATA" magic, DataBlock::
CopyHeaderToFile (IO/UVF/DataBlock.cpp:72-82) for every block, and for the TOC
block the ExtendedOctree right after its block header, written by
ExtendedOctree::WriteHeader (IO/UVF/ExtendedOctree/ExtendedOctree.cpp:499-549)
in the writing machine's (little-endian) order, the bricks after the TOC.
Test infrastructure: the bricks come from the oracle's conversion.
"""
import struct
import zlib

import numpy as np

BS_EMPTY, BS_KEY_VALUE_PAIRS, BS_TOC_BLOCK = 0, 4, 9
CT_NONE, CT_ZLIB, CT_LZMA = 0, 1, 2


def _block_header(block_id, semantics, payload_len, big_endian, last):
    e = ">" if big_endian else "<"
    head = struct.pack(e + "Q", len(block_id)) + block_id + struct.pack(e + "QQ", semantics, 0)
    nxt = 0 if last else len(head) + 8 + payload_len
    return head + struct.pack(e + "Q", nxt)


def octree_bytes(ds, aspect=(1.0, 1.0, 1.0), codec=lambda i: CT_NONE, order=None, atlas=None, version=2):
    """The ExtendedOctree of an oracle dataset: header, TOC (ExtendedOctree index
    order = the dataset's TOC order) and the bricks, stored in `order` (a
    permutation of the brick indices; default TOC order)."""
    n = ds.brick_count
    bricks = [np.ascontiguousarray(ds.brick_data(i)).tobytes() for i in range(n)]
    comp = [codec(i) for i in range(n)]
    payload = [zlib.compress(b, 6) if c == CT_ZLIB else b for b, c in zip(bricks, comp)]
    nx, ny, nz = ds.dims
    head = struct.pack("<IQB", 0 if ds.bpv == 1 else 1, 1, 0)
    head += struct.pack("<QQQ", nx, ny, nz) + struct.pack("<ddd", *aspect)
    head += struct.pack("<QQQ", *ds.brick) + struct.pack("<I", ds.overlap)
    tail_len = 4 + 8 + (4 if version > 1 else 0)       # version, size, compression level
    toc_len = n * (8 + 8 + 4 + 8 + 4 + 4)
    data_at = len(head) + tail_len + toc_len
    order = list(range(n)) if order is None else list(order)
    offsets, pos = [0] * n, data_at
    for i in order:
        offsets[i] = pos
        pos += len(payload[i])
    head += struct.pack("<IQ", version, pos)           # m_iSize: header and bricks
    if version > 1:
        head += struct.pack("<I", 6)                    # compression level
    toc = b"".join(struct.pack("<QQIQII", offsets[i], len(payload[i]), comp[i], len(payload[i]),
                               *(atlas(i) if atlas else (0, 0))) for i in range(n))
    return head + toc + b"".join(payload[i] for i in order)


def write_uvf(path, ds, big_endian=False, extra_block=True, **octree_kw):
    """A version-5 UVF file: global header, an optional key/value block before
    the TOC block (the reader has to step over it), then the TOC block."""
    e = ">" if big_endian else "<"
    out = bytearray(b"UVF-DATA")
    out += struct.pack("B", 1 if big_endian else 0)
    out += struct.pack(e + "QQQ", 5, 0, 0)                # version, checksum semantics, checksum length
    out += struct.pack(e + "Q", 0)                        # offset to the first data block
    if extra_block:
        kv = b"\x00" * 40                                 # an opaque payload the reader skips
        out += _block_header(b"key/value pairs", BS_KEY_VALUE_PAIRS, len(kv), big_endian, last=False) + kv
    octree = octree_bytes(ds, **octree_kw)
    out += _block_header(b"Table of Contents Raster Data", BS_TOC_BLOCK, len(octree), big_endian, last=True)
    out += octree
    with open(path, "wb") as f:
        f.write(out)
    return len(out)
