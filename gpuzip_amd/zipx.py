"""ZIP entries decoded in batches on the GPU: the reference extractor's batched GPU path
(src/extract.rs:458-477 take_gpu_batch, 511-596 gpu_thread, 598-628 writer; src/gpu.rs:345-408
Decoder::decode over nvCOMP) on this library's gi_inflate_many.

The reference hands batches of its smallest Deflate entries to nvCOMP: each batch is copied
into a pinned slot, decoded, and every entry's size and CRC-32 checked against the central
directory before it is written. Here a batch is one gi_inflate_many call: the entries' spans in
the archive go to the library as they lie (it uploads them), each is decoded with the others of
its batch segment in one set of launches, and the CRC-32 comes back from the GPU with the
output. What the reference then gives its CPU threads (entries a batch could not take, or got
wrong) is reported here as failed, never decoded on the host: this module is the GPU path only.

The directory is read as the reference's src/zip.rs reads it (end record, ZIP64 end record and
locator, the central directory's sizes and offsets with their ZIP64 extra field, each entry's data
after its local header's name and extra field)."""
import struct
import zlib

STORED, DEFLATED = 0, 8
ALIGN = 16                 # an entry's room in a slot, rounded up (src/gpu.rs:19-24)
MAX_BATCH_ENTRIES = 65536  # entries a batch holds at most (src/gpu.rs:20)


class ZipError(Exception):
    pass


class Entry:
    __slots__ = ("name", "method", "flags", "crc", "csize", "size", "header_offset", "data_offset")

    def __repr__(self):
        return f"Entry({self.name!r}, method {self.method}, {self.csize} -> {self.size})"


def align(n):
    return (n + ALIGN - 1) // ALIGN * ALIGN


def _end_record(data):
    """(entries, central directory size, its offset) from the end of central directory record,
    through the ZIP64 end record when a field is saturated"""
    lo = max(0, len(data) - (65535 + 22))
    at = data.rfind(b"PK\x05\x06", lo)
    if at < 0 or at + 22 > len(data):
        raise ZipError("no end of central directory record")
    _, _, _, _, n, cd_size, cd_off, _ = struct.unpack_from("<IHHHHIIH", data, at)
    if n == 0xFFFF or cd_size == 0xFFFFFFFF or cd_off == 0xFFFFFFFF:
        loc = at - 20
        if loc < 0 or data[loc:loc + 4] != b"PK\x06\x07":
            raise ZipError("ZIP64 sizes without a ZIP64 end record locator")
        (e64,) = struct.unpack_from("<Q", data, loc + 8)
        if e64 + 56 > len(data) or data[e64:e64 + 4] != b"PK\x06\x06":
            raise ZipError("the ZIP64 end record is not where its locator says")
        n, cd_size, cd_off = struct.unpack_from("<QQQ", data, e64 + 32)
    return n, cd_size, cd_off


def read_entries(data):
    """The archive's entries, in central directory order, with their data offsets"""
    data = memoryview(data).tobytes() if not isinstance(data, (bytes, bytearray)) else data
    n, cd_size, at = _end_record(data)
    if at + cd_size > len(data):
        raise ZipError("the central directory lies past the end of the archive")
    out = []
    for _ in range(n):
        if data[at:at + 4] != b"PK\x01\x02":
            raise ZipError(f"no central directory header at byte {at}")
        (flags, method, crc, csize, size, nlen, xlen, clen, hoff) = struct.unpack_from(
            "<HHxxxxIIIHHH8xI", data, at + 8)
        name = data[at + 46:at + 46 + nlen]
        extra = data[at + 46 + nlen:at + 46 + nlen + xlen]
        # ZIP64: the saturated fields, in this order, in extra field 0x0001
        x = 0
        while x + 4 <= len(extra):
            tag, ln = struct.unpack_from("<HH", extra, x)
            if tag == 1:
                vals, y = [], x + 4
                for sat in (size == 0xFFFFFFFF, csize == 0xFFFFFFFF, hoff == 0xFFFFFFFF):
                    if sat:
                        vals.append(struct.unpack_from("<Q", extra, y)[0])
                        y += 8
                    else:
                        vals.append(None)
                size = vals[0] if vals[0] is not None else size
                csize = vals[1] if vals[1] is not None else csize
                hoff = vals[2] if vals[2] is not None else hoff
            x += 4 + ln
        e = Entry()
        e.name = name.decode("utf-8" if flags & 0x800 else "cp437").replace("\\", "/")
        e.method, e.flags, e.crc, e.csize, e.size, e.header_offset = method, flags, crc, csize, size, hoff
        if hoff + 30 > len(data) or data[hoff:hoff + 4] != b"PK\x03\x04":
            raise ZipError(f"{e.name}: no local header at byte {hoff}")
        ln_, lx_ = struct.unpack_from("<HH", data, hoff + 26)
        e.data_offset = hoff + 30 + ln_ + lx_
        if e.data_offset + csize > len(data):
            raise ZipError(f"{e.name}: its data runs past the end of the archive")
        out.append(e)
        at += 46 + nlen + xlen + clen
    return out


def gpu_batches(entries, slot_bytes, max_entry):
    """The Deflate entries in batches, smallest first, as take_gpu_batch forms them
    (extract.rs:458-477): a batch takes entries while it has fewer than MAX_BATCH_ENTRIES, each is
    at most max_entry, and the entries' aligned sizes and compressed sizes each fit slot_bytes.
    Entries over max_entry (or larger than a slot) are left out: [indices of a batch], [left out]"""
    pending = sorted((i for i, e in enumerate(entries) if e.method == DEFLATED), key=lambda i: entries[i].size)
    batches, left = [], []
    k = 0
    while k < len(pending):
        batch, used_u, used_c = [], 0, 0
        while k < len(pending):
            e = entries[pending[k]]
            if e.size > max_entry or align(e.size) > slot_bytes or align(e.csize) > slot_bytes:
                left.extend(pending[k:])  # smallest first: every later one is as large
                k = len(pending)
                break
            u, c = align(e.size), align(e.csize)
            if len(batch) == MAX_BATCH_ENTRIES or used_u + u > slot_bytes or used_c + c > slot_bytes:
                break
            batch.append(pending[k])
            used_u += u
            used_c += c
            k += 1
        if batch:
            batches.append(batch)
    return batches, left


def extract(data, gpu, slot_bytes=1 << 30, max_entry=1 << 30):
    """Every entry of the archive `data`: Deflate entries decoded on `gpu` (a gpuinflate.Gpu) in
    gi_inflate_many batches, stored ones taken as they lie; each output's size and CRC-32 checked
    against the central directory (the Deflate ones with the CRC-32 the GPU computed). Returns
    ({name: bytes}, [(name, why) for entries not extracted], {"batches": n, "gpu_entries": n,
    "gpu_bytes": n})."""
    entries = read_entries(data)
    files, failed = {}, []
    batches, left = gpu_batches(entries, slot_bytes, max_entry)
    gpu_entries = gpu_bytes = 0
    for batch in batches:
        spans = [(entries[i].data_offset, entries[i].csize, align(entries[i].size)) for i in batch]
        res, outs, _ = gpu.inflate_many(data, spans)
        for i, r, out in zip(batch, res, outs):
            e = entries[i]
            if r.status != 0:
                failed.append((e.name, f"GPU status {r.status}"))
            elif r.out_size != e.size or r.crc32 != e.crc:
                failed.append((e.name, "size or CRC-32 differs from the central directory"))
            else:
                files[e.name] = out
                gpu_entries += 1
                gpu_bytes += e.size
    for i in left:
        failed.append((entries[i].name, "larger than a GPU batch takes"))
    for e in entries:
        if e.method == STORED:
            out = data[e.data_offset:e.data_offset + e.size]
            if len(out) != e.size or zlib.crc32(out) != e.crc:
                failed.append((e.name, "size or CRC-32 differs from the central directory"))
            else:
                files[e.name] = bytes(out)
        elif e.method != DEFLATED:
            failed.append((e.name, f"method {e.method} is not read here"))
    return files, failed, {"batches": len(batches), "gpu_entries": gpu_entries, "gpu_bytes": gpu_bytes}
