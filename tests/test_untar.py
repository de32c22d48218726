"""A .tar.gz extracted while the GPU decodes it (gpuzip_amd/untar.py, the reference's tar pipe,
src/gz.rs:266-289, 622-822): member paths kept inside the output folder, the pipe's pieces read
in order, and on the GPU every file equal to tarfile's, a path with ".." skipped, a damaged tar
stream inside a good gzip member a corrupt archive."""
import gzip
import io
import os
import queue
import tarfile
import threading

import pytest

from gpuzip_amd import datagen, untar


def test_tar_path():
    assert untar.tar_path("a/b/c.txt") == "a/b/c.txt"
    assert untar.tar_path("/abs/x") == "abs/x"
    assert untar.tar_path("./a/./b") == "a/b"
    assert untar.tar_path("a/../../etc/passwd") is None
    assert untar.tar_path("..") is None
    assert untar.tar_path("a\\b") == "a/b"


def test_pipe_reads_pieces_in_order():
    q = queue.Queue(2)
    p = untar._Pipe(q)
    pieces = [bytes([i]) * (1000 + 37 * i) for i in range(40)]

    def produce():
        for b in pieces:
            q.put(b)
        q.put(None)
    t = threading.Thread(target=produce)
    t.start()
    got = []
    for n in (1, 999, 5000, 3, 100000, 7, 250000):
        got.append(p.read(n))
    got.append(p.read())
    t.join()
    assert b"".join(got) == b"".join(pieces)
    assert p.read(10) == b""


def make_tar(files, extra=()):
    buf = io.BytesIO()
    with tarfile.open(fileobj=buf, mode="w", format=tarfile.PAX_FORMAT) as t:
        for name, data in files.items():
            if data is None:
                ti = tarfile.TarInfo(name)
                ti.type = tarfile.DIRTYPE
                t.addfile(ti)
                continue
            ti = tarfile.TarInfo(name)
            ti.size = len(data)
            t.addfile(ti, io.BytesIO(data))
        for name, data in extra:
            ti = tarfile.TarInfo(name)
            ti.size = len(data)
            t.addfile(ti, io.BytesIO(data))
    return buf.getvalue()


@pytest.mark.gpu
def test_tar_gz_extracted_as_it_decodes(tmp_path):
    from gpuzip_amd import gz
    from gpuzip_amd.gpuinflate import Gpu
    big = datagen.mixed(40 << 20, seed=3, run_max=1 << 20).tobytes()
    files = {"top": None, "top/sub": None, "top/big.log": big, "top/sub/empty": b"",
             "top/sub/small.txt": b"hello tar pipe\n" * 1000, "readme": big[:123457]}
    tar = make_tar(files, extra=[("../escape.txt", b"nope")])
    member = gzip.compress(tar, 6)
    with Gpu() as g:
        u = untar.untar_gz(g, member, str(tmp_path / "out"))
        assert u.error is None and [n for n, _ in u.failed] == ["../escape.txt"]
        assert u.files == 4 and u.bytes == sum(len(v) for v in files.values() if v)
        for name, data in files.items():
            p = tmp_path / "out" / name
            if data is None:
                assert p.is_dir()
            else:
                assert p.read_bytes() == data, name
        assert not (tmp_path / "escape.txt").exists()
        assert not any(n.endswith(".gpuzip-part") for n in os.listdir(tmp_path / "out" / "top"))
        # listing only, two members
        u2 = untar.untar_gz(g, gzip.compress(tar[:len(tar) // 2], 1) + gzip.compress(tar[len(tar) // 2:], 9))
        assert [n for n, _ in u2.listed] == list(files) + ["../escape.txt"]
        # a good gzip member holding a damaged tar stream
        bad = bytearray(tar)
        bad[512 * 3 + 100:512 * 3 + 200] = b"\xff" * 100
        with pytest.raises(gz.Corrupt):
            untar.untar_gz(g, gzip.compress(bytes(bad[: 512 * 4]), 6))


def test_reader_on_a_fed_pipe(tmp_path):
    """the tar side alone, fed by a thread as the GPU's sink would feed it: files written inside
    the output folder, an escaping path skipped, a stream cut inside a member damaged"""
    big = datagen.mixed(1 << 20, seed=3, run_max=1 << 18).tobytes()
    files = {"top": None, "top/big.log": big, "top/sub/empty": b"", "readme": big[:1000]}
    tar = make_tar(files, extra=[("../escape.txt", b"nope")])

    def run(data, root):
        q = queue.Queue(untar.PIPE_DEPTH)
        u = untar.Untarred()
        t = threading.Thread(target=untar._untar, args=(untar._Pipe(q), root, u))
        t.start()
        for a in range(0, len(data), 4096):
            q.put(data[a:a + 4096])
        q.put(None)
        t.join()
        return u
    u = run(tar, str(tmp_path))
    assert u.error is None and u.files == 3 and [n for n, _ in u.failed] == ["../escape.txt"]
    assert (tmp_path / "top" / "big.log").read_bytes() == big and (tmp_path / "readme").read_bytes() == big[:1000]
    u = run(tar[: 512 * 4], None)
    assert u.error is not None and "damaged" in u.error
