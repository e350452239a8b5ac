"""Zarr v2 directory stores, read and written without the zarr package.

The visibility store that the reference's ``convert_ms``/``write_vis`` write
(cngi/conversion/convert_ms.py:156-184, cngi/dio/write_vis.py:26-91) is a
directory of zarr v2 groups written by ``xarray.Dataset.to_zarr``: per group a
``.zgroup``, the dataset attributes in ``.zattrs``, optionally the
consolidated ``.zmetadata``, and per variable a directory with ``.zarray``
(shape, chunks, dtype, order, fill value, compressor, filters), ``.zattrs``
(``_ARRAY_DIMENSIONS`` and the CF encoding attributes) and one file per chunk
named by its chunk indices joined with the dimension separator.  These
formats are public: the zarr v2 storage specification, the numcodecs codec
configurations (``{"id": ...}``), the Blosc 1 frame layout and the xarray
zarr/CF encoding conventions; this module restates what the visibility
layout uses.

Codecs: none, ``zlib``, ``gzip``, ``bz2``, ``lzma`` (Python's standard
library), ``zstd`` (a zstd frame, through pyarrow), and ``blosc`` frames whose
inner codec is zstd, lz4, zlib or snappy, with byte shuffle or none (the
reference's default is Blosc zstd level 2 without shuffle).  Blosc's own
``blosclz`` codec and bit shuffle are refused with an error naming them.
Filters: ``vlen-utf8`` (variable-length strings).  Stores written here use the
``zstd`` codec by default.
"""
import bz2
import gzip
import json
import lzma
import math
import os
import struct
import zlib

import numpy as np

_FMT_TO_CODEC = {1: "lz4_raw", 2: "snappy", 3: "zlib", 4: "zstd"}   # Blosc flag bits 5-7
_NAME_TO_FMT = {"lz4": 1, "lz4hc": 1, "snappy": 2, "zlib": 3, "zstd": 4, "blosclz": 0}


class ZarrFormatError(ValueError):
    """A store uses a layout or codec this reader does not handle."""


# ---------------------------------------------------------------------------
# codecs
# ---------------------------------------------------------------------------
def _arrow_codec(name, level=None):
    import pyarrow as pa
    if level is None:
        return pa.Codec(name)
    return pa.Codec(name, compression_level=level)


def _zstd_content_size(frame):
    """Frame_Content_Size of a zstd frame (RFC 8878, 3.1.1.1), or None."""
    if len(frame) < 6 or struct.unpack_from("<I", frame, 0)[0] != 0xFD2FB528:
        raise ZarrFormatError("not a zstd frame")
    fhd = frame[4]
    fcs_flag, single = fhd >> 6, (fhd >> 5) & 1
    pos = 5 + (0 if single else 1) + (0, 1, 2, 4)[fhd & 3]
    size = (1 if single else 0, 2, 4, 8)[fcs_flag]
    if size == 0:
        return None
    v = int.from_bytes(frame[pos:pos + size], "little")
    return v + 256 if size == 2 else v


def _zstd_decompress(data, size=None):
    """A zstd frame's content, of the size its header states (else `size`)."""
    fcs = _zstd_content_size(data)
    if fcs is None and size is None:
        raise ZarrFormatError("zstd frame without a content size")
    return _arrow_codec("zstd").decompress(data, decompressed_size=fcs if fcs is not None
                                           else size, asbytes=True)


def _blosc_decompress(buf):
    """Decode one Blosc 1 frame: a 16-byte header (version, codec version,
    flags, type size, uncompressed bytes, block size, compressed bytes), then
    the byte offset of every block, each block holding one stream (or one per
    byte of the type when split) as a 4-byte size and the compressed bytes
    (a size equal to the stream's length marks stored bytes)."""
    if len(buf) < 16:
        raise ZarrFormatError("short Blosc frame")
    flags, typesize = buf[2], buf[3]
    nbytes, blocksize, cbytes = struct.unpack_from("<III", buf, 4)
    if flags & 0x2:   # memcpyed: stored as is after the header
        return bytes(buf[16:16 + nbytes])
    if flags & 0x4:
        raise ZarrFormatError("Blosc bit shuffle is not supported")
    if flags & 0x8:
        raise ZarrFormatError("Blosc delta coding is not supported")
    fmt = flags >> 5
    if fmt not in _FMT_TO_CODEC:
        raise ZarrFormatError("Blosc codec 'blosclz' is not available here")
    codec = _FMT_TO_CODEC[fmt]
    dont_split = bool(flags & 0x10)
    nblocks = max(1, math.ceil(nbytes / blocksize)) if nbytes else 0
    starts = struct.unpack_from("<%di" % nblocks, buf, 16)
    out = bytearray(nbytes)
    for b in range(nblocks):
        bsize = min(blocksize, nbytes - b * blocksize)
        leftover = bsize < blocksize
        split = (not dont_split and typesize <= 16 and bsize // max(typesize, 1) >= 128
                 and not leftover)
        nsplits = typesize if split else 1
        neb = bsize // nsplits
        pos = starts[b]
        parts = []
        for _ in range(nsplits):
            csize = struct.unpack_from("<i", buf, pos)[0]
            pos += 4
            chunk = bytes(buf[pos:pos + csize])
            pos += csize
            if csize == neb:
                parts.append(chunk)
            elif codec == "zlib":
                parts.append(zlib.decompress(chunk))
            else:
                parts.append(_arrow_codec(codec).decompress(chunk, decompressed_size=neb,
                                                            asbytes=True))
        block = b"".join(parts)
        if (flags & 0x1) and typesize > 1:   # byte shuffle within the block
            n = bsize // typesize
            body = np.frombuffer(block[:n * typesize], dtype=np.uint8)
            block = body.reshape(typesize, n).T.tobytes() + block[n * typesize:]
        out[b * blocksize:b * blocksize + bsize] = block
    return bytes(out)


def _blosc_compress(data, cname="zstd", clevel=2, shuffle=0, typesize=1, blocksize=0):
    """A Blosc 1 frame of `data` (no split streams, one block unless
    blocksize is given), for the codecs _blosc_decompress reads."""
    fmt = _NAME_TO_FMT[cname]
    if fmt == 0:
        raise ZarrFormatError("Blosc codec 'blosclz' is not available here")
    nbytes = len(data)
    bs = blocksize or max(nbytes, 1)
    nblocks = max(1, math.ceil(nbytes / bs)) if nbytes else 0
    flags = 0x10 | (fmt << 5) | (0x1 if shuffle == 1 and typesize > 1 else 0)
    blocks = []
    for b in range(nblocks):
        blk = data[b * bs:(b + 1) * bs]
        if flags & 0x1:
            n = len(blk) // typesize
            body = np.frombuffer(blk[:n * typesize], dtype=np.uint8)
            blk = body.reshape(n, typesize).T.tobytes() + blk[n * typesize:]
        if fmt == 3:
            c = zlib.compress(blk, clevel)
        else:
            c = _arrow_codec(_FMT_TO_CODEC[fmt], clevel if fmt == 4 else None).compress(
                blk, asbytes=True)
        if len(c) >= len(blk):
            c = blk   # stored
        blocks.append(struct.pack("<i", len(c)) + c)
    head = 16 + 4 * nblocks
    starts, pos = [], head
    for blk in blocks:
        starts.append(pos)
        pos += len(blk)
    header = struct.pack("<BBBBIII", 2, 1, flags, typesize, nbytes, bs, pos)
    return header + struct.pack("<%di" % nblocks, *starts) + b"".join(blocks)


def decode_chunk(raw, compressor, size=None):
    """Chunk bytes as stored -> uncompressed bytes (size: expected length)."""
    if compressor is None:
        return raw
    cid = compressor.get("id")
    if cid == "zlib":
        return zlib.decompress(raw)
    if cid == "gzip":
        return gzip.decompress(raw)
    if cid == "bz2":
        return bz2.decompress(raw)
    if cid == "lzma":
        return lzma.decompress(raw)
    if cid == "zstd":
        return _zstd_decompress(raw, size)
    if cid == "blosc":
        return _blosc_decompress(raw)
    raise ZarrFormatError(f"compressor {cid!r} is not supported")


def encode_chunk(data, compressor, typesize=1):
    if compressor is None:
        return data
    cid = compressor.get("id")
    level = compressor.get("level", compressor.get("clevel"))
    if cid == "zlib":
        return zlib.compress(data, 1 if level is None else level)
    if cid == "gzip":
        return gzip.compress(data, 1 if level is None else level)
    if cid == "bz2":
        return bz2.compress(data, 1 if level is None else level)
    if cid == "lzma":
        return lzma.compress(data)
    if cid == "zstd":
        return _arrow_codec("zstd", level).compress(data, asbytes=True)
    if cid == "blosc":
        return _blosc_compress(data, compressor.get("cname", "zstd"),
                               compressor.get("clevel", 2), compressor.get("shuffle", 0),
                               typesize, compressor.get("blocksize", 0))
    raise ZarrFormatError(f"compressor {cid!r} is not supported")


def _vlen_utf8_decode(buf, n_items):
    """numcodecs VLenUTF8: item count (uint32), then per item its byte length
    (uint32) and UTF-8 bytes."""
    n = struct.unpack_from("<I", buf, 0)[0]
    if n != n_items:
        raise ZarrFormatError("vlen-utf8 item count does not match the chunk")
    out = np.empty(n, dtype=object)
    pos = 4
    for i in range(n):
        ln = struct.unpack_from("<I", buf, pos)[0]
        pos += 4
        out[i] = bytes(buf[pos:pos + ln]).decode("utf-8")
        pos += ln
    return out


def _vlen_utf8_encode(arr):
    parts = [struct.pack("<I", arr.size)]
    for s in arr.ravel(order="C"):
        b = ("" if s is None else str(s)).encode("utf-8")
        parts.append(struct.pack("<I", len(b)) + b)
    return b"".join(parts)


# ---------------------------------------------------------------------------
# arrays
# ---------------------------------------------------------------------------
def _json(path):
    with open(path) as f:
        return json.load(f)


def _fill_value(meta, dtype):
    fv = meta.get("fill_value")
    if fv is None or dtype == object:
        return None
    if isinstance(fv, str):
        if fv in ("NaN", "Infinity", "-Infinity"):
            return {"NaN": np.nan, "Infinity": np.inf, "-Infinity": -np.inf}[fv]
        if dtype.kind in "SUV":
            return fv
        raise ZarrFormatError(f"fill value {fv!r} not supported")
    return fv


def read_array(path, meta=None, attrs=None):
    """(numpy array, attributes) of the zarr v2 array at `path`."""
    meta = _json(os.path.join(path, ".zarray")) if meta is None else meta
    if attrs is None:
        zattrs = os.path.join(path, ".zattrs")
        attrs = _json(zattrs) if os.path.exists(zattrs) else {}
    if meta.get("zarr_format") != 2:
        raise ZarrFormatError("only zarr v2 arrays are read")
    filters = meta.get("filters") or []
    vlen = any(f.get("id") == "vlen-utf8" for f in filters)
    if any(f.get("id") != "vlen-utf8" for f in filters):
        raise ZarrFormatError(f"filters {filters} not supported")
    descr = meta["dtype"]
    dtype = np.dtype(object) if (vlen or descr == "|O") else np.dtype(
        [tuple(x) for x in descr] if isinstance(descr, list) else descr)
    shape = tuple(meta["shape"])
    chunks = tuple(meta["chunks"])
    order = meta.get("order", "C")
    sep = meta.get("dimension_separator", ".")
    fv = _fill_value(meta, dtype)
    out = np.empty(shape, dtype=dtype)
    if dtype != object:
        out[...] = 0 if fv is None else fv
    if len(shape) == 0:
        grid = [()]
    else:
        grid = np.ndindex(*[max(1, math.ceil(s / c)) for s, c in zip(shape, chunks)])
    for idx in grid:
        name = sep.join(str(i) for i in idx) if len(shape) else "0"
        f = os.path.join(path, name)
        if not os.path.exists(f):
            continue   # a missing chunk holds the fill value
        with open(f, "rb") as fh:
            raw = fh.read()
        cshape = chunks if len(shape) else ()
        n_items = int(np.prod(cshape)) if len(shape) else 1
        data = decode_chunk(raw, meta.get("compressor"),
                            None if vlen else n_items * dtype.itemsize)
        if vlen or dtype == object:
            block = _vlen_utf8_decode(data, n_items).reshape(cshape, order=order)
        else:
            block = np.frombuffer(data, dtype=dtype, count=n_items).reshape(cshape, order=order)
        sl = tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, chunks, shape))
        out[sl] = block[tuple(slice(0, x.stop - x.start) for x in sl)]
    return out, attrs


def write_array(path, arr, chunks=None, compressor=None, attrs=None):
    """Write `arr` as a zarr v2 array (C order, '.' separator)."""
    arr = np.asarray(arr)
    os.makedirs(path, exist_ok=True)
    shape = arr.shape
    if chunks is None:
        chunks = shape
    chunks = tuple(max(1, min(int(c) if c and c > 0 else s, s) if s else 1)
                   for c, s in zip(chunks, shape))
    vlen = arr.dtype.kind in "OU"
    if arr.dtype.kind == "U":
        arr = arr.astype(object)
    meta = {"zarr_format": 2, "shape": list(shape), "chunks": list(chunks),
            "dtype": "|O" if vlen else arr.dtype.str, "order": "C",
            "fill_value": None if vlen else 0 if arr.dtype.kind not in "fc" else "NaN",
            "compressor": compressor,
            "filters": [{"id": "vlen-utf8"}] if vlen else None,
            "dimension_separator": "."}
    if not vlen and arr.dtype.kind == "b":
        meta["fill_value"] = False
    with open(os.path.join(path, ".zarray"), "w") as f:
        json.dump(meta, f, indent=2)
    with open(os.path.join(path, ".zattrs"), "w") as f:
        json.dump(attrs or {}, f, indent=2)
    grid = np.ndindex(*[max(1, math.ceil(s / c)) for s, c in zip(shape, chunks)]) \
        if len(shape) else [()]
    for idx in grid:
        sl = tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, chunks, shape))
        block = arr[sl]
        if block.shape != chunks:   # edge chunks are stored at full chunk shape
            full = np.empty(chunks, dtype=arr.dtype)
            if vlen:
                full[...] = ""
            else:
                full[...] = 0
            full[tuple(slice(0, n) for n in block.shape)] = block
            block = full
        raw = _vlen_utf8_encode(block) if vlen else np.ascontiguousarray(block).tobytes()
        name = ".".join(str(i) for i in idx) if len(shape) else "0"
        with open(os.path.join(path, name), "wb") as f:
            f.write(encode_chunk(raw, compressor, 1 if vlen else arr.dtype.itemsize))
    return meta


# ---------------------------------------------------------------------------
# groups (xarray datasets)
# ---------------------------------------------------------------------------
_TIME_UNITS = {"days": 86400e9, "hours": 3600e9, "minutes": 60e9, "seconds": 1e9,
               "milliseconds": 1e6, "microseconds": 1e3, "nanoseconds": 1.0}


def _decode_cf(arr, attrs):
    """CF time decoding as xarray applies it: '<unit> since <date>'."""
    units = attrs.get("units")
    if isinstance(units, str) and " since " in units and arr.dtype.kind in "iuf":
        unit, ref = units.split(" since ", 1)
        scale = _TIME_UNITS.get(unit.strip())
        if scale is not None:
            ref = np.datetime64(ref.strip().replace(" ", "T").rstrip("Z"), "ns")
            if arr.dtype.kind in "iu":   # exact in integers
                ns = arr.astype(np.int64) * np.int64(scale)
            else:
                ns = np.round(arr.astype(np.float64) * scale).astype(np.int64)
            out = ref + ns.astype("timedelta64[ns]")
            if "_FillValue" in attrs and arr.dtype.kind == "f":
                out = np.where(np.isnan(arr), np.datetime64("NaT"), out)
            return out, {k: v for k, v in attrs.items() if k not in ("units", "calendar")}
    if "_FillValue" in attrs and arr.dtype.kind == "f":
        attrs = {k: v for k, v in attrs.items() if k != "_FillValue"}
    return arr, attrs


def read_group(path, consolidated=True):
    """(variables {name: (dims, array, attrs)}, group attributes) of a zarr v2
    group written by xarray.  Coordinates are the variables named like a
    dimension or listed in a variable's 'coordinates' attribute."""
    zmeta = os.path.join(path, ".zmetadata")
    metas = None
    if consolidated and os.path.exists(zmeta):
        metas = _json(zmeta).get("metadata", {})
    if metas is None:
        metas = {}
        for name in [".zgroup", ".zattrs"]:
            f = os.path.join(path, name)
            if os.path.exists(f):
                metas[name] = _json(f)
        for d in sorted(os.listdir(path)):
            for name in [".zarray", ".zattrs"]:
                f = os.path.join(path, d, name)
                if os.path.exists(f):
                    metas[f"{d}/{name}"] = _json(f)
    if ".zgroup" not in metas:
        raise ZarrFormatError(f"{path} is not a zarr group")
    gattrs = metas.get(".zattrs", {})
    variables = {}
    for key, meta in metas.items():
        if not key.endswith("/.zarray") or key.count("/") != 1:
            continue
        name = key.split("/")[0]
        attrs = dict(metas.get(f"{name}/.zattrs", {}))
        arr, _ = read_array(os.path.join(path, name), meta, attrs)
        dims = tuple(attrs.pop("_ARRAY_DIMENSIONS", [f"dim_{i}" for i in range(arr.ndim)]))
        arr, attrs = _decode_cf(arr, attrs)
        variables[name] = (dims, arr, attrs)
    return variables, gattrs


def write_group(path, variables, attrs, compressor=None, chunks=None, consolidated=True):
    """Write a zarr v2 group: variables {name: (dims, array, attrs)}; chunks
    {dim: size} (the whole axis for dims not named)."""
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, ".zgroup"), "w") as f:
        json.dump({"zarr_format": 2}, f)
    with open(os.path.join(path, ".zattrs"), "w") as f:
        json.dump(_jsonable(attrs), f, indent=2)
    metas = {".zgroup": {"zarr_format": 2}, ".zattrs": _jsonable(attrs)}
    for name, (dims, arr, vattrs) in variables.items():
        arr = np.asarray(arr)
        a = dict(_jsonable(vattrs or {}))
        if arr.dtype.kind == "M":   # datetime64 -> integer ns since the epoch (CF)
            arr = arr.astype("datetime64[ns]").astype(np.int64)
            a["units"] = "nanoseconds since 1970-01-01"
            a["calendar"] = "proleptic_gregorian"
        a["_ARRAY_DIMENSIONS"] = list(dims)
        ch = [((chunks or {}).get(d) or s) for d, s in zip(dims, arr.shape)]
        meta = write_array(os.path.join(path, name), arr, ch, compressor, a)
        metas[f"{name}/.zarray"] = meta
        metas[f"{name}/.zattrs"] = a
    if consolidated:
        with open(os.path.join(path, ".zmetadata"), "w") as f:
            json.dump({"metadata": metas, "zarr_consolidated_format": 1}, f, indent=2)


def update_group_attrs(path, extra):
    f = os.path.join(path, ".zattrs")
    attrs = _json(f) if os.path.exists(f) else {}
    attrs.update(_jsonable(extra))
    with open(f, "w") as fh:
        json.dump(attrs, fh, indent=2)


def consolidate(path):
    """Write .zmetadata: every .zgroup/.zattrs/.zarray of the group and its
    arrays under their relative keys."""
    metas = {}
    for name in (".zgroup", ".zattrs"):
        f = os.path.join(path, name)
        if os.path.exists(f):
            metas[name] = _json(f)
    for d in sorted(os.listdir(path)):
        for name in (".zarray", ".zattrs"):
            f = os.path.join(path, d, name)
            if os.path.exists(f):
                metas[f"{d}/{name}"] = _json(f)
    with open(os.path.join(path, ".zmetadata"), "w") as fh:
        json.dump({"metadata": metas, "zarr_consolidated_format": 1}, fh, indent=2)


def _jsonable(v):
    if isinstance(v, dict):
        return {str(k): _jsonable(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [_jsonable(x) for x in v]
    if isinstance(v, np.ndarray):
        return _jsonable(v.tolist())
    if isinstance(v, np.generic):
        return v.item()
    return v
