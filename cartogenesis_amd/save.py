"""The ocean's four fields in the reference's save format, format 28, through ``include/cgsave.h``.

A save written by the reference (WorldCodec.write, cartography/src/commonMain/kotlin/com/
cartogenesis/cartography/WorldCodec.kt:431-477) carries the ocean the stage produced as four binary32
sections, ocean.velocityX, ocean.velocityY, ocean.temperature and ocean.anomaly, one value per cell of
the map grid (WorldSections.kt:184-187). This module reads them out of a save and writes a save whose
ocean is replaced, every other byte of the world carried as it was, so a world whose ocean was solved
here reopens in the reference unchanged (desktop/src/test/kotlin/com/cartogenesis/desktop/
GpuOceanTest.kt:122-135). The work is the library's (no GPU is used); a save that the reference
would refuse is refused here with the reference's reason, as ``SaveRefused``.
"""
from __future__ import annotations

import ctypes
import mmap
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import CG_ERR_NOMEM, CG_OK, lib

CG_SAVE_NOT_A_SAVE, CG_SAVE_INCOMPLETE, CG_SAVE_DAMAGED = -20, -21, -22
CG_SAVE_WRONG_VERSION, CG_SAVE_TOO_LARGE, CG_SAVE_CANNOT_EXPAND = -23, -24, -25
CG_SAVE_FORMAT_VERSION = 28
CG_SAVE_CHUNK_BYTES = 1 << 20
CG_SAVE_SECTIONS = 57
SECTION_TYPES = {1: "F32", 2: "I32", 3: "U8", 4: "UTF8"}

# SaveProblem (cartography/.../SaveRefusal.kt:11-45), by the library's status codes.
PROBLEMS = {
    CG_SAVE_NOT_A_SAVE: "NOT_A_SAVE",
    CG_SAVE_INCOMPLETE: "INCOMPLETE",
    CG_SAVE_DAMAGED: "DAMAGED",
    CG_SAVE_WRONG_VERSION: "WRONG_VERSION",
    CG_SAVE_TOO_LARGE: "TOO_LARGE",
    CG_SAVE_CANNOT_EXPAND: "CANNOT_EXPAND",
}

OCEAN_SECTIONS = ("ocean.velocityX", "ocean.velocityY", "ocean.temperature", "ocean.anomaly")


class SectionInfo(ctypes.Structure):
    """cg_section_info: a directory entry (WorldSections.kt:54-67)."""
    _fields_ = [("name", ctypes.c_char * 48), ("type", ctypes.c_int), ("count", ctypes.c_int),
                ("bytes", ctypes.c_int64), ("offset", ctypes.c_int64)]


class SaveRefused(RuntimeError):
    """WorldFormatException: a save that does not open, with its SaveProblem and the reason."""

    def __init__(self, problem: str, detail: str):
        super().__init__(f"{problem}: {detail}" if detail else problem)
        self.problem = problem
        self.detail = detail


@dataclass(frozen=True)
class OceanFields:
    """The four arrays, binary32, (down, across) row-major: row 0 is the northern edge."""
    velocity_x: np.ndarray
    velocity_y: np.ndarray
    temperature: np.ndarray
    anomaly: np.ndarray

    def arrays(self) -> tuple[np.ndarray, ...]:
        return self.velocity_x, self.velocity_y, self.temperature, self.anomaly

    @staticmethod
    def empty(across: int, down: int) -> "OceanFields":
        return OceanFields(*(np.empty((down, across), np.float32) for _ in range(4)))


@dataclass(frozen=True)
class SaveGrid:
    across: int
    down: int
    payload_bytes: int
    largest_rewrite: int


def _bytes_of(save) -> np.ndarray:
    if isinstance(save, np.ndarray):
        arr = save.reshape(-1).view(np.uint8)
    else:
        arr = np.frombuffer(save, dtype=np.uint8)
    if not arr.flags.c_contiguous:
        arr = np.ascontiguousarray(arr)
    return arr


def _fields_in(fields: OceanFields, cells: int) -> list[np.ndarray]:
    out = []
    for name, a in zip(OCEAN_SECTIONS, fields.arrays()):
        a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
        if a.size != cells:
            raise ValueError(f"{name} holds {a.size} cells where the save's grid has {cells}")
        out.append(a)
    return out


def _raise(status: int, reason: ctypes.Array, what: str) -> None:
    if status == CG_OK:
        return
    detail = reason.value.decode("utf-8", "replace")
    if status in PROBLEMS:
        raise SaveRefused(PROBLEMS[status], detail)
    if status == CG_ERR_NOMEM:
        raise MemoryError(detail or what)
    raise ValueError(f"{what}: {detail or 'bad arguments'} ({status})")


def crc32(data, continuing: int = 0) -> int:
    """Crc32.of (PngWriter.kt:378-407)."""
    b = _bytes_of(data)
    return int(lib().cg_save_crc32(b.ctypes.data if b.size else None, b.size, continuing))


def chunk_checksum(header_checksum: int, index: int, raw) -> int:
    """chunkChecksum (WorldSections.kt:496-499)."""
    b = _bytes_of(raw)
    return int(lib().cg_save_chunk_checksum(header_checksum, index, b.ctypes.data if b.size else None, b.size))


def directory(cells: int, lists_bytes: int) -> tuple[list[dict], int]:
    """WorldSections.directory and payloadBytes (WorldSections.kt:251-265): the entries as the
    header's JSON spells them, and the expanded payload's length."""
    entries = (SectionInfo * CG_SAVE_SECTIONS)()
    total = ctypes.c_int64()
    n = lib().cg_save_directory(cells, lists_bytes, ctypes.cast(entries, ctypes.c_void_p), CG_SAVE_SECTIONS,
                                ctypes.byref(total))
    if n != CG_SAVE_SECTIONS:
        raise ValueError(f"no directory for {cells} cells and {lists_bytes} bytes of lists")
    return [{"name": e.name.decode(), "type": SECTION_TYPES[e.type], "count": e.count, "bytes": e.bytes,
             "offset": e.offset} for e in entries], total.value


def ocean_records(fields: OceanFields) -> bytes:
    """The four ocean records as they stand in the expanded payload of a save of this grid."""
    cells = fields.velocity_x.size
    arrays = _fields_in(fields, cells)
    need = ctypes.c_size_t()
    size = sum(20 + len(n) + 4 * cells for n in OCEAN_SECTIONS)
    out = np.empty(size, np.uint8)
    status = lib().cg_ocean_records_write(cells, *arrays, out.ctypes.data, out.size, ctypes.byref(need))
    if status != CG_OK:
        bad = [n for n, a in zip(OCEAN_SECTIONS, arrays) if not np.isfinite(a).all()]
        raise ValueError(f"the ocean records cannot be written: {bad or 'bad arguments'} ({status})")
    return out[:need.value].tobytes()


def ocean_from_records(records, across: int, down: int) -> OceanFields:
    """The reverse of ``ocean_records``, every record held to its directory entry."""
    b = _bytes_of(records)
    f = OceanFields.empty(across, down)
    reason = ctypes.create_string_buffer(512)
    status = lib().cg_ocean_records_read(b.ctypes.data if b.size else None, b.size, across * down,
                                         *(a.reshape(-1) for a in f.arrays()), reason, len(reason))
    _raise(status, reason, "ocean records")
    return f


def inspect(save) -> SaveGrid:
    """The grid and sizes of a save, from its prefix and header (WorldCodec.decodeHeader)."""
    b = _bytes_of(save)
    across, down = ctypes.c_int(), ctypes.c_int()
    payload, largest = ctypes.c_int64(), ctypes.c_int64()
    reason = ctypes.create_string_buffer(512)
    status = lib().cg_save_inspect(b.ctypes.data if b.size else None, b.size, ctypes.byref(across),
                                   ctypes.byref(down), ctypes.byref(payload), ctypes.byref(largest), reason,
                                   len(reason))
    _raise(status, reason, "inspect")
    return SaveGrid(across.value, down.value, payload.value, largest.value)


def read_ocean(save) -> OceanFields:
    """The ocean of a whole save, checked from its first byte to its last (WorldCodec.read,
    WorldCodec.kt:525-546). ``save`` is bytes, a NumPy byte array or a memory map."""
    b = _bytes_of(save)
    grid = inspect(b)
    f = OceanFields.empty(grid.across, grid.down)
    reason = ctypes.create_string_buffer(512)
    status = lib().cg_save_read_ocean(b.ctypes.data, b.size, *(a.reshape(-1) for a in f.arrays()), reason,
                                      len(reason))
    _raise(status, reason, "read")
    return f


def replace_ocean(save, fields: OceanFields, store: bool = False) -> np.ndarray:
    """The save with its ocean replaced, as a NumPy byte array: the same header and chunks, the four
    arrays' elements written over, each chunk gzip-compressed again when the header names gzip and
    ``store`` is false (and compressing makes it shorter), stored otherwise."""
    b = _bytes_of(save)
    grid = inspect(b)
    arrays = _fields_in(fields, grid.across * grid.down)
    out = np.empty(grid.largest_rewrite, np.uint8)
    written = ctypes.c_size_t()
    reason = ctypes.create_string_buffer(512)
    status = lib().cg_save_replace_ocean(b.ctypes.data, b.size, *arrays, int(store), out.ctypes.data, out.size,
                                         ctypes.byref(written), reason, len(reason))
    _raise(status, reason, "replace")
    return out[:written.value]


def open_save(path) -> np.ndarray:
    """A save file mapped read-only: read a chunk at a time by the functions above, never loaded whole."""
    with open(Path(path), "rb") as fh:
        mm = mmap.mmap(fh.fileno(), 0, access=mmap.ACCESS_READ)
    return np.frombuffer(mm, dtype=np.uint8)


def save_with_ocean(source, fields: OceanFields, target, store: bool = False) -> int:
    """Writes the save at ``source`` with its ocean replaced to ``target``; returns its length."""
    out = replace_ocean(open_save(source), fields, store)
    out.tofile(str(target))
    return int(out.size)


__all__ = ["OCEAN_SECTIONS", "OceanFields", "SaveGrid", "SaveRefused", "chunk_checksum", "crc32", "directory",
           "inspect", "ocean_from_records", "ocean_records", "open_save", "read_ocean", "replace_ocean",
           "save_with_ocean"]
