"""A .tar.gz extracted while the GPU decodes it: the reference host's tar pipe (src/gz.rs:266-289
Pipe, 622-780 untar, 806-822 its use in run) over this library's member driver.

The decoded bytes leave the GPU through gz_decode's sink (libgzhost.so: every member decoded on
the GPU, its CRC-32 and size checked against its trailer) in pieces of at most PIPE_CHUNK, and go
through a bounded queue of 8 pieces to a reader thread that walks the tar stream as it arrives,
so the archive is never whole in host memory. The tar reader is Python's own (the reference uses
the tar crate); what it writes follows the reference's rules: regular files and folders only,
each member's path kept inside the output folder (a path with a ".." component is failed and
skipped, a leading "/" or "./" dropped), a file written under a temporary name and renamed once
whole. A damaged tar stream after good gzip members is a corrupt archive, as in the reference
(gz.rs:818-821); when the sink stops because the reader failed, the reader's error is the one
reported."""
import os
import queue
import tarfile
import threading

from . import gz

PIPE_CHUNK = 4 << 20  # the tar reader gets the output in pieces this big (src/gz.rs:20)
PIPE_DEPTH = 8        # pieces in flight (src/gz.rs:807)


class _Pipe:
    """the decoded bytes, as a file the tar reader reads (src/gz.rs:266-289)"""

    def __init__(self, q):
        self.q = q
        self.cur = b""
        self.pos = 0
        self.done = False

    def read(self, n=-1):
        out = []
        while n != 0:
            if self.pos == len(self.cur):
                if self.done:
                    break
                item = self.q.get()
                if item is None:
                    self.done = True
                    break
                self.cur, self.pos = item, 0
            take = len(self.cur) - self.pos if n < 0 else min(n, len(self.cur) - self.pos)
            out.append(self.cur[self.pos:self.pos + take])
            self.pos += take
            if n > 0:
                n -= take
        return b"".join(out)


def tar_path(name):
    """the member's path inside the output folder, or None when it would leave it (gz.rs:303-315)"""
    parts = []
    for c in name.replace("\\", "/").split("/"):
        if c in ("", "."):
            continue
        if c == "..":
            return None
        parts.append(c)
    return "/".join(parts)


class Untarred:
    def __init__(self):
        self.files = 0
        self.bytes = 0
        self.failed = []   # (name, why)
        self.listed = []   # (name, size) of every member, in order
        self.error = None  # the tar stream is damaged


def _untar(pipe, root, u):
    try:
        ar = tarfile.open(fileobj=pipe, mode="r|")
    except tarfile.TarError as e:
        u.error = f"not a tar archive: {e}"
        return
    try:
        for m in ar:
            u.listed.append((m.name, m.size if m.isfile() else None))
            if root is None or not (m.isfile() or m.isdir()):
                if m.isfile():
                    u.files += 1
                    u.bytes += m.size
                continue
            rel = tar_path(m.name)
            if rel is None:
                u.failed.append((m.name, "path leaves the output folder; skipped"))
                continue
            dst = os.path.join(root, rel) if rel else root
            if m.isdir():
                os.makedirs(dst, exist_ok=True)
                continue
            os.makedirs(os.path.dirname(dst) or root, exist_ok=True)
            src = ar.extractfile(m)
            tmp = dst + ".gpuzip-part"
            with open(tmp, "wb") as f:
                while True:
                    b = src.read(PIPE_CHUNK)
                    if not b:
                        break
                    f.write(b)
            os.replace(tmp, dst)
            u.files += 1
            u.bytes += m.size
        # the padding after the end-of-archive blocks, so that the decoder's checks run to the end
        while pipe.read(PIPE_CHUNK):
            pass
    except (tarfile.TarError, OSError) as e:
        u.error = f"tar archive is damaged: {e}"


def untar_gz(gpu, data, root=None):
    """Extracts (root given) or lists the .tar.gz `data` on gpu (a gpuinflate.Gpu) as it decodes.
    Returns Untarred; raises gz.Corrupt for a damaged gzip or tar stream, gz.Retry for a GPU
    failure (the reference then decodes on the CPU)."""
    q = queue.Queue(PIPE_DEPTH)
    pipe = _Pipe(q)
    u = Untarred()
    if root is not None:
        os.makedirs(root, exist_ok=True)
    reader = threading.Thread(target=_untar, args=(pipe, root, u))
    reader.start()

    def sink(b):
        # a reader that stopped early (a damaged tar stream) no longer takes pieces: stop the decode
        for a in range(0, len(b), PIPE_CHUNK):
            while True:
                if not reader.is_alive():
                    return False
                try:
                    q.put(b[a:a + PIPE_CHUNK], timeout=0.1)
                    break
                except queue.Full:
                    continue
        return True

    err = None
    try:
        gz.decode(gpu, data, sink)
    except gz.Stopped:
        err = "stopped"
    except (gz.Corrupt, gz.Retry) as e:
        err = e
    finally:
        while reader.is_alive():
            try:
                q.put(None, timeout=0.1)
                break
            except queue.Full:
                continue
        reader.join()
    if u.error is not None and (err is None or err == "stopped"):
        raise gz.Corrupt(u.error)
    if isinstance(err, Exception):
        raise err
    return u
