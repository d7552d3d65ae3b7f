"""CPU restatement of the reference's PLY loader (src/io/ply.ts).

TEST INFRASTRUCTURE ONLY: the checker for the device PLY ingestion
(dl_scene_create_from_ply); never imported by the product.

What is restated, function by function:
  parse_header      <- parseHeader, src/io/ply.ts:40-91
  load_ply_splats   <- loadPlySplats, src/io/ply.ts:120-239
  js_hypot          <- Math.hypot as the JS engines implement it (the
                       max-scaled, Kahan-compensated sum of squares used by
                       V8's math.tq and JavaScriptCore's MathObject.cpp), which
                       ply.ts:220 calls
  splat_bounds      <- splatBounds, src/io/ply.ts:242-255

JS numbers are f64: SH_C0 * f and w / n are evaluated in f64 and rounded to
f32 when stored into the Float32Arrays; numpy f64 arithmetic is the same IEEE
arithmetic.
"""
from __future__ import annotations

import re
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

SH_C0 = 0.28209479177387814
HEADER_MAX = 65536
_SIZES = {"float": 4, "uchar": 1, "int": 4}

# ECMAScript WhiteSpace + LineTerminator code points (String.prototype.trim)
_JS_WS = {0x09, 0x0B, 0x0C, 0x20, 0xA0, 0xFEFF, 0x0A, 0x0D, 0x2028, 0x2029, 0x1680, 0x2000, 0x2001, 0x2002,
          0x2003, 0x2004, 0x2005, 0x2006, 0x2007, 0x2008, 0x2009, 0x200A, 0x202F, 0x205F, 0x3000}


class PlyError(ValueError):
    pass


@dataclass
class PlyHeader:
    count: int
    properties: List[Tuple[str, str]]  # (type, name)
    data_offset: int


def _js_trim(s: str) -> str:
    a, b = 0, len(s)
    while a < b and ord(s[a]) in _JS_WS:
        a += 1
    while b > a and ord(s[b - 1]) in _JS_WS:
        b -= 1
    return s[a:b]


def _js_split_ws(s: str) -> List[str]:
    """line.split(/\\s+/) (JS \\s = WhiteSpace + LineTerminator)."""
    out, cur = [], []
    i = 0
    while i < len(s):
        if ord(s[i]) in _JS_WS:
            out.append("".join(cur))
            cur = []
            while i < len(s) and ord(s[i]) in _JS_WS:
                i += 1
            continue
        cur.append(s[i])
        i += 1
    out.append("".join(cur))
    return out


def _utf16_len(s: str) -> int:
    return sum(2 if ord(c) > 0xFFFF else 1 for c in s)


def _js_parse_int(s: str) -> float:
    """parseInt(s, 10): leading whitespace, optional sign, decimal digits."""
    s = _js_trim(s)
    m = re.match(r"[+-]?[0-9]+", s)
    if not m:
        return float("nan")
    return float(int(m.group(0)))


def parse_header(data: bytes) -> PlyHeader:
    """parseHeader (ply.ts:40-91). The data offset is the UTF-16 index of
    'end_header\\n' in the decoded text plus 11, as the reference computes
    it (equal to the byte offset for an ASCII header)."""
    text = bytes(data[:min(HEADER_MAX, len(data))]).decode("utf-8", errors="replace")
    end = text.find("end_header\n")
    if end < 0:
        raise PlyError("PLY: end_header not found in first 64KB")
    data_offset = _utf16_len(text[:end]) + len("end_header\n")
    lines = [t for t in (_js_trim(l) for l in text[:end].split("\n")) if t]
    if not lines or lines[0] != "ply":
        raise PlyError("PLY: missing magic")
    count = 0.0
    binary, little = False, True
    props: List[Tuple[str, str]] = []
    in_vertex = False
    for line in lines[1:]:
        if line.startswith("format "):
            binary = "binary" in line
            little = "little_endian" in line
            continue
        if line.startswith("element vertex "):
            count = _js_parse_int(line[len("element vertex "):])
            in_vertex = True
            continue
        if line.startswith("element "):
            in_vertex = False
            continue
        if in_vertex and line.startswith("property "):
            parts = _js_split_ws(line)
            t = parts[1] if len(parts) > 1 else "undefined"
            name = parts[2] if len(parts) > 2 else "undefined"
            if t not in ("float", "uchar", "int"):
                raise PlyError(f'PLY: unsupported property type "{t}"')
            props.append((t, name))
    if not binary:
        raise PlyError("PLY: ASCII PLY not supported (binary only)")
    if not little:
        raise PlyError("PLY: only little-endian supported")
    if not (count > 0):
        # count <= 0; a NaN count (no digits) is rejected here as well, where
        # the reference would go on with count = NaN
        raise PlyError("PLY: no vertices")
    return PlyHeader(int(count), props, data_offset)


def js_hypot(*cols: np.ndarray) -> np.ndarray:
    """Math.hypot of f64 arrays, element-wise, as the engines compute it:
    +inf if any argument is infinite; NaN if any is NaN; 0 if all are 0;
    else sqrt(Kahan sum of (x / max)^2) * max."""
    a = [np.abs(np.asarray(c, np.float64)) for c in cols]
    mx = np.zeros_like(a[0])
    for x in a:
        mx = np.where(x > mx, x, mx)
    anynan = np.zeros(mx.shape, bool)
    anyinf = np.zeros(mx.shape, bool)
    for x in a:
        anynan |= np.isnan(x)
        anyinf |= np.isinf(x)
    safe = np.where((mx == 0) | anynan | anyinf, 1.0, mx)
    s = np.zeros_like(mx)
    comp = np.zeros_like(mx)
    with np.errstate(invalid="ignore", over="ignore"):
        for x in a:
            q = x / safe
            summand = q * q - comp
            prelim = s + summand
            comp = (prelim - s) - summand
            s = prelim
        out = np.sqrt(s) * safe
    out = np.where(mx == 0, 0.0, out)
    out = np.where(anynan, np.nan, out)
    return np.where(anyinf, np.inf, out)


@dataclass
class PlySplats:
    count: int
    positions: np.ndarray
    rotations: np.ndarray
    log_scales: np.ndarray
    opacity_logits: np.ndarray
    colors_dc: np.ndarray
    colors_sh: Optional[np.ndarray]
    sh_bands: int
    coeffs_per_channel: int


def load_ply_splats(data: bytes, flip_y: bool = True) -> PlySplats:
    """loadPlySplats (ply.ts:120-239)."""
    hdr = parse_header(data)
    idx = {}
    for p, (_, name) in enumerate(hdr.properties):
        idx[name] = p
    if not all(t == "float" for t, _ in hdr.properties):
        raise PlyError("PLY: mixed property types not yet supported")
    P = len(hdr.properties)
    total = hdr.count * P * 4
    if hdr.data_offset + total > len(data):
        raise PlyError(f"PLY: data section truncated (need {total} bytes, have {len(data) - hdr.data_offset})")
    flat = np.frombuffer(bytes(data[hdr.data_offset:hdr.data_offset + total]), dtype="<f4").reshape(hdr.count, P)

    def need(n: str) -> int:
        if n not in idx:
            raise PlyError(f'PLY: missing property "{n}"')
        return idx[n]

    ix, iy, iz = need("x"), need("y"), need("z")
    isx, isy, isz = need("scale_0"), need("scale_1"), need("scale_2")
    id0, id1, id2 = need("f_dc_0"), need("f_dc_1"), need("f_dc_2")
    iop = need("opacity")
    ir = [need("rot_0"), need("rot_1"), need("rot_2"), need("rot_3")]
    k_rest = 0
    while f"f_rest_{k_rest}" in idx:
        k_rest += 1
    cpc = k_rest // 3
    bands = 3 if cpc >= 15 else 2 if cpc >= 8 else 1 if cpc >= 3 else 0
    n = hdr.count
    f64 = flat.astype(np.float64)
    ysign = -1.0 if flip_y else 1.0
    pos = np.stack([flat[:, ix], (ysign * f64[:, iy]).astype(np.float32), flat[:, iz]], axis=1)
    ls = np.stack([flat[:, isx], flat[:, isy], flat[:, isz]], axis=1)
    op = flat[:, iop].copy()
    dc = (SH_C0 * f64[:, [id0, id1, id2]]).astype(np.float32)
    w, qx, qy, qz = (f64[:, i] for i in ir)
    nrm = js_hypot(w, qx, qy, qz)
    nrm = np.where((nrm == 0) | np.isnan(nrm), 1.0, nrm)  # `|| 1`
    f = -1.0 if flip_y else 1.0
    with np.errstate(invalid="ignore"):
        rot = np.stack([w / nrm, (f * qx) / nrm, qy / nrm, (f * qz) / nrm], axis=1).astype(np.float32)
    sh = None
    if bands > 0:
        sh = flat[:, [idx[f"f_rest_{k}"] for k in range(3 * cpc)]].copy()
    return PlySplats(n, pos.reshape(-1), rot.reshape(-1), ls.reshape(-1).copy(), op, dc.reshape(-1),
                     None if sh is None else sh.reshape(-1), bands, cpc)


def splat_bounds(positions: np.ndarray) -> Tuple[List[float], List[float]]:
    """splatBounds (ply.ts:242-255): a strict `<` / `>` scan from +-Infinity,
    so NaN coordinates never win and among equal values (+0 and -0) the
    first in index order stays."""
    p = np.asarray(positions, np.float32).reshape(-1, 3).astype(np.float64)
    lo, hi = [], []
    for c in range(3):
        v = p[:, c]
        v = v[~np.isnan(v)]
        if not v.size:
            lo.append(float("inf"))
            hi.append(float("-inf"))
            continue
        mn, mx = v.min(), v.max()
        lo.append(float(v[np.flatnonzero(v == mn)[0]]))
        hi.append(float(v[np.flatnonzero(v == mx)[0]]))
    return lo, hi


def write_ply(props: List[Tuple[str, str]], rows: np.ndarray, fmt: str = "binary_little_endian 1.0",
              extra_header: str = "", pre_vertex: str = "", post_vertex: str = "") -> bytes:
    """A binary PLY with the given vertex properties (all written as the
    rows' little-endian f32 values); test fixtures only."""
    head = ["ply", f"format {fmt}"]
    if extra_header:
        head.append(extra_header)
    if pre_vertex:
        head.append(pre_vertex)
    head.append(f"element vertex {rows.shape[0]}")
    head += [f"property {t} {n}" for t, n in props]
    if post_vertex:
        head.append(post_vertex)
    head.append("end_header")
    return ("\n".join(head) + "\n").encode() + np.ascontiguousarray(rows, dtype="<f4").tobytes()
