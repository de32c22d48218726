"""ZIP entries in GPU batches (gpuzip_amd/zipx.py): the central directory read as Python's zipfile
reads it (ZIP64 included), batches formed by the reference's rule (src/extract.rs:458-477), and on
the GPU every entry equal to zipfile's bytes, a damaged entry failed alone."""
import io
import random
import zipfile
import zlib

import pytest

from gpuzip_amd import datagen, zipx


def archive(n=300, seed=4, zip64=False, stored_every=7):
    rng = random.Random(seed)
    pool = datagen.mixed(8 << 20, seed=seed, run_max=256 << 10).tobytes()
    buf = io.BytesIO()
    want = {}
    with zipfile.ZipFile(buf, "w") as z:
        for i in range(n):
            size = 0 if i % 50 == 3 else int(2 ** rng.uniform(0, 20))
            a = rng.randrange(0, len(pool) - size) if size < len(pool) else 0
            d = pool[a:a + size]
            name = f"dir{i % 5}/file_{i:04d}.txt"
            method = zipfile.ZIP_STORED if i % stored_every == 0 else zipfile.ZIP_DEFLATED
            info = zipfile.ZipInfo(name)
            info.compress_type = method
            info._compresslevel = (1, 6, 9)[i % 3]
            with z.open(info, "w", force_zip64=zip64) as f:
                f.write(d)
            want[name] = d
    return buf.getvalue(), want


@pytest.mark.parametrize("zip64", [False, True])
def test_directory_as_zipfile_reads_it(zip64, monkeypatch):
    if zip64:
        # zipfile writes ZIP64 fields past these limits: lowered, small entries get ZIP64 sizes
        # and offsets in the central directory, and the end record goes through ZIP64 too
        monkeypatch.setattr(zipfile, "ZIP64_LIMIT", 1 << 12)
        monkeypatch.setattr(zipfile, "ZIP_FILECOUNT_LIMIT", 50)
    data, want = archive(120, seed=zip64 + 1, zip64=zip64)
    if zip64:
        assert b"PK\x06\x06" in data[-200:] and b"PK\x06\x07" in data[-100:]
    entries = zipx.read_entries(data)
    infos = zipfile.ZipFile(io.BytesIO(data)).infolist()
    assert [e.name for e in entries] == [i.filename for i in infos]
    for e, i in zip(entries, infos):
        assert (e.method, e.crc, e.csize, e.size, e.header_offset) == (
            i.compress_type, i.CRC, i.compress_size, i.file_size, i.header_offset)
        raw = data[e.data_offset:e.data_offset + e.csize]
        got = raw if e.method == zipx.STORED else zlib.decompress(raw, -15)
        assert got == want[e.name]


def test_batches_follow_the_reference_rule():
    data, _ = archive(300, seed=9)
    entries = zipx.read_entries(data)
    slot, max_entry = 1 << 20, 256 << 10
    batches, left = zipx.gpu_batches(entries, slot, max_entry)
    deflate = [i for i, e in enumerate(entries) if e.method == zipx.DEFLATED]
    flat = [i for b in batches for i in b]
    assert sorted(flat + left) == sorted(deflate)
    sizes = [entries[i].size for i in flat]
    assert sizes == sorted(sizes)  # smallest first
    assert all(entries[i].size > max_entry for i in left)
    for b in batches:
        assert sum(zipx.align(entries[i].size) for i in b) <= slot
        assert sum(zipx.align(entries[i].csize) for i in b) <= slot
    # each batch but the last stopped because the next entry would not fit
    for b, nxt in zip(batches, batches[1:]):
        e = entries[nxt[0]]
        assert (sum(zipx.align(entries[i].size) for i in b) + zipx.align(e.size) > slot or
                sum(zipx.align(entries[i].csize) for i in b) + zipx.align(e.csize) > slot)


def test_not_a_zip():
    with pytest.raises(zipx.ZipError):
        zipx.read_entries(b"PK\x03\x04 but nothing else")


@pytest.mark.gpu
def test_extract_on_the_gpu():
    from gpuzip_amd.gpuinflate import Gpu
    data, want = archive(300, seed=5)
    with Gpu(segment_out_mb=256) as g:
        files, failed, st = zipx.extract(data, g, slot_bytes=4 << 20)
        assert not failed, failed[:3]
        assert files == want
        assert st["batches"] > 2 and st["gpu_entries"] > 200
        # a damaged entry fails alone (its CRC or its data), the rest as before
        entries = zipx.read_entries(data)
        k = max((i for i, e in enumerate(entries) if e.method == zipx.DEFLATED), key=lambda i: entries[i].csize)
        bad = bytearray(data)
        e = entries[k]
        for j in range(e.data_offset + e.csize // 2, e.data_offset + e.csize // 2 + 16):
            bad[j] ^= 0xA5
        files2, failed2, _ = zipx.extract(bytes(bad), g, slot_bytes=4 << 20)
        assert [n for n, _ in failed2] == [e.name]
        assert {n: v for n, v in files2.items()} == {n: v for n, v in want.items() if n != e.name}
