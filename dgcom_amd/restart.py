"""The reference's 2D restart file (`DGRST2D`, src/io/restart.rs:1-67,405-736):
a run's state between two steps, so a run can stop and continue bit for bit.

Layout (little-endian): the magic `DGRST2D` and the format version (1); the
element count and nodes per element (u64 each); the domain fingerprint (u64);
the model time (f64); the section count (u64), then per section its name
(u32 length + UTF-8 bytes), its kind (u8: 0 f64 values, 1 UTF-8 text), its
length (u64: values or bytes) and its payload; last the FNV-1a (64 bit)
checksum of every byte before it. Sections: `state.h`, `state.hu`, `state.hv`,
the driver's arrays `extra.<name>`, texts `text.<name>`, and `metadata`
(`key=value` lines) last. A write goes to `<path>.tmp` first and is renamed
over the target, so an interrupted write keeps the previous file.

The domain fingerprint is FNV-1a over the mesh's vertex count, every vertex's
x and y bits, the element count, every element's four vertex indices and the
bed's length and bits (restart.rs:660-697): a restart resumes only on the
domain it was written from.
"""
from __future__ import annotations

import dataclasses
import os
import struct

import numpy as np

MAGIC_2D = b"DGRST2D"
MAGIC_3D = b"DGRST3D"
VERSION = 1
KIND_VALUES, KIND_TEXT = 0, 1
FNV_OFFSET = 0xCBF29CE484222325
FNV_PRIME = 0x100000001B3
_MASK = (1 << 64) - 1
SWE_FIELDS = ("h", "hu", "hv")


class RestartError(Exception):
    """A restart that cannot be read or does not belong to the run (the
    reference's RestartError: Io, Format, Domain)."""


def fnv1a(data: bytes, h: int = FNV_OFFSET) -> int:
    """FNV-1a (64 bit) over `data`, continuing from `h` (the library's
    dg_fnv1a: a byte-serial recurrence, run natively)."""
    from . import swe2d
    L = swe2d.load()
    buf = np.frombuffer(data, dtype=np.uint8)
    return int(L.dg_fnv1a(buf.ctypes.data_as(swe2d._u8p), buf.size, h))


def mesh_vertices(mesh):
    """The mesh's shared vertices [n][2] and elements' vertex indices [K][4]:
    the ones it was built with (`structured_mesh` keeps its grid's, numbered as
    the reference's structured_quads, mesh2d.rs:216-243), else the corners'
    distinct points in order of first appearance."""
    if getattr(mesh, "vertices", None) is not None and getattr(mesh, "quads", None) is not None:
        return np.asarray(mesh.vertices, dtype=np.float64), np.asarray(mesh.quads, dtype=np.uint64)
    pts = np.asarray(mesh.corners, dtype=np.float64).reshape(-1, 2)
    _, first, inverse = np.unique(pts, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first)
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    return pts[first[order]], rank[inverse.ravel()].reshape(-1, 4).astype(np.uint64)


def domain_fingerprint(mesh, bed) -> int:
    """FNV-1a over the mesh's vertices and elements and the bed's bits
    (domain_fingerprint_2d, restart.rs:690-697)."""
    vertices, quads = mesh_vertices(mesh)
    bed = np.ascontiguousarray(bed, dtype=np.float64)
    words = np.concatenate([
        np.array([vertices.shape[0]], dtype=np.uint64),
        np.ascontiguousarray(vertices, dtype=np.float64).view(np.uint64).ravel(),
        np.array([quads.shape[0]], dtype=np.uint64),
        quads.ravel(),
        np.array([bed.size], dtype=np.uint64),
        bed.view(np.uint64),
    ]).astype("<u8")
    return fnv1a(words.tobytes())


class _Hashing:
    """A binary writer that hashes what passes through it."""

    def __init__(self, f):
        self.f = f
        self.h = FNV_OFFSET

    def write(self, b: bytes):
        self.h = fnv1a(b, self.h)
        self.f.write(b)


@dataclasses.dataclass
class Restart2D:
    """A 2D run's state between two steps (Restart2D, restart.rs:187-235)."""
    time: float
    domain: int
    h: np.ndarray
    hu: np.ndarray
    hv: np.ndarray
    n_elements: int
    n_nodes: int
    metadata: list = dataclasses.field(default_factory=list)    # [(key, value)]
    text: list = dataclasses.field(default_factory=list)        # [(name, text)]
    extra: list = dataclasses.field(default_factory=list)       # [(name, f64 array)]

    def write(self, path: str):
        """Write to `path` through `<path>.tmp`, then rename over `path`."""
        sections = [(f"state.{n}", np.ascontiguousarray(v, dtype="<f8")) for n, v in
                    zip(SWE_FIELDS, (self.h, self.hu, self.hv))]
        sections += [(f"extra.{n}", np.ascontiguousarray(v, dtype="<f8")) for n, v in self.extra]
        for k, v in self.metadata:
            if not k or any(c in k for c in "=\n\r") or any(c in v for c in "\n\r"):
                raise ValueError(f"a restart metadata pair must have no '=' in its key and no newlines: {k!r}")
        meta = "".join(f"{k}={v}\n" for k, v in self.metadata).encode()
        tmp = str(path) + ".tmp"
        with open(tmp, "wb") as f:
            out = _Hashing(f)
            out.write(MAGIC_2D + bytes([VERSION]))
            out.write(struct.pack("<QQQd", self.n_elements, self.n_nodes, self.domain, self.time))
            out.write(struct.pack("<Q", len(sections) + len(self.text) + 1))
            for name, vals in sections:
                nb = name.encode()
                out.write(struct.pack("<I", len(nb)) + nb + bytes([KIND_VALUES]) + struct.pack("<Q", vals.size))
                out.write(vals.tobytes())
            for name, txt in self.text:
                nb = f"text.{name}".encode()
                tb = txt.encode()
                out.write(struct.pack("<I", len(nb)) + nb + bytes([KIND_TEXT]) + struct.pack("<Q", len(tb)))
                out.write(tb)
            nb = b"metadata"
            out.write(struct.pack("<I", len(nb)) + nb + bytes([KIND_TEXT]) + struct.pack("<Q", len(meta)))
            out.write(meta)
            f.write(struct.pack("<Q", out.h))
            f.flush()
            os.fsync(f.fileno())
        os.replace(tmp, path)

    @staticmethod
    def read(path: str) -> "Restart2D":
        """Read a restart written by `write`, checking its magic, version,
        sections and checksum (restart.rs:518-660)."""
        with open(path, "rb") as f:
            data = f.read()
        if len(data) < 8 + 8:
            raise RestartError(f"{path}: truncated")
        body, (checksum,) = data[:-8], struct.unpack("<Q", data[-8:])
        if data[:7] != MAGIC_2D:
            raise RestartError("a 3D restart, not a 2D one" if data[:7] == MAGIC_3D else "bad magic number")
        if data[7] != VERSION:
            raise RestartError(f"restart format version {data[7]}, this reads {VERSION}")
        if fnv1a(body) != checksum:
            raise RestartError(f"{path}: checksum mismatch (the file is damaged or truncated)")
        pos = 8
        n_elements, n_nodes, domain, time = struct.unpack_from("<QQQd", body, pos)
        pos += 32
        (n_sections,) = struct.unpack_from("<Q", body, pos)
        pos += 8
        values, text, meta = {}, [], ""
        for _ in range(n_sections):
            (ln,) = struct.unpack_from("<I", body, pos)
            pos += 4
            name = body[pos:pos + ln].decode()
            pos += ln
            kind = body[pos]
            pos += 1
            (n,) = struct.unpack_from("<Q", body, pos)
            pos += 8
            if kind == KIND_VALUES:
                values[name] = np.frombuffer(body, dtype="<f8", count=n, offset=pos).astype(np.float64)
                pos += 8 * n
            elif kind == KIND_TEXT:
                txt = body[pos:pos + n].decode()
                pos += n
                if name == "metadata":
                    meta = txt
                elif name.startswith("text."):
                    text.append((name[5:], txt))
            else:
                raise RestartError(f"section {name!r} of unknown kind {kind}")
        if pos != len(body):
            raise RestartError(f"{len(body) - pos} trailing bytes before the checksum")
        nt = n_elements * n_nodes
        state = []
        for fld in SWE_FIELDS:
            v = values.get(f"state.{fld}")
            if v is None or v.size != nt:
                raise RestartError(f"section state.{fld} missing or not {nt} values")
            state.append(v)
        metadata = [tuple(line.split("=", 1)) for line in meta.splitlines() if line]
        extra = [(k[6:], v) for k, v in values.items() if k.startswith("extra.")]
        return Restart2D(time, domain, *state, n_elements, n_nodes, metadata, text, extra)

    def check_domain(self, mesh, bed, order: int):
        """SWEPhysics2D::check_restart: the restart is of this domain."""
        nn = (order + 1) ** 2
        if (self.n_elements, self.n_nodes) != (mesh.n_elements, nn):
            raise RestartError(f"the restart has {self.n_elements} elements of {self.n_nodes} nodes, "
                               f"the run {mesh.n_elements} of {nn}")
        fp = domain_fingerprint(mesh, bed)
        if fp != self.domain:
            raise RestartError(f"the restart is of another domain (fingerprint {self.domain:#x}, the run's {fp:#x})")
