"""Schedule autotuner for the batched decoder (gi_batch_*, include/gpuinflate.h).

The reference has no counterpart: its GPU decoder's schedule is fixed in constants
(gpuinflate/gpuinflate.cu:59-75, one job per 4 KiB chunk at most, probes every 16 KiB with a
4 KiB lead, 256 MiB segments). Here the schedule is a run-time choice made by measurement on
the machine and the workload at hand:

- ``workers``: decoder contexts in the batch, 1 or 2 searched (each with its own two HIP
  streams; one plans on the host while another's kernels run). The library takes up to 4, but a
  process has four hardware queues, so a third context's streams share them and its finder
  graphs queue behind the others' decodes;
- ``segment_out_mb``: a batch segment's output at most (0: from free memory, 4 GiB at most),
  held at 0 by the search;
- ``chunk_kib``: the search stride (one speculative job per chunk of input at most);
- ``probe_spacing_kbit`` and ``probe_lead_bits``: where stretches with no block start get
  probes, and how far before its point a probe starts decoding;
- ``lanes_per_group``: how many lanes decode one interval together (16, 32 or 64: a quarter,
  a half or a whole wavefront per interval; narrower groups split an interval less finely and
  hold more code tables per compute unit's LDS);
- ``persistent_per_cu``: the group decoder as one persistent launch of this many workgroups a
  compute unit taking work from a device queue, or (0) a workgroup per queue item.

Results do not depend on the schedule (every job's bytes, size, CRC-32 and ``in_used`` are
those of decoding it alone), so a candidate whose results differ from the expected ones is
rejected, never chosen. The search is coordinate descent from the library's defaults: each
field in turn takes each of its values with the others held, a change is kept when it beats
the best so far by ``min_gain``, and passes repeat until one changes nothing.
"""

import json
import os
import statistics
import time

# the values each field may take; the first pass starts from DEFAULT
SPACE = {
    "workers": [1, 2],
    "chunk_kib": [4, 8, 16],
    "probe_spacing_kbit": [256, 512, 1024],
    "probe_lead_bits": [4096, 8192],
    "lanes_per_group": [16, 32, 64],
    "persistent_per_cu": [0, 10],
}
DEFAULT = {"workers": 2, "segment_out_mb": 0, "chunk_kib": 4, "probe_spacing_kbit": 1024, "probe_lead_bits": 4096,
           "lanes_per_group": 64, "persistent_per_cu": 0}
# the fields that go into gi_config_tuned (the rest are gi_batch_create's and gi_config's)
TUNED_FIELDS = ("chunk_kib", "probe_spacing_kbit", "probe_lead_bits", "lanes_per_group", "persistent_per_cu")


def _key(cfg):
    return tuple(sorted(cfg.items()))


def coordinate_search(measure, space=None, start=None, passes=3, min_gain=0.02, log=None, confirm=True):
    """Coordinate descent over `space` (field -> values) from `start`, maximising measure(cfg).

    measure returns a score (higher is better) or None for a rejected candidate. Each distinct
    candidate is measured once during the search. With `confirm`, a winner other than the
    start is measured again beside the start (start, winner, start: one noisy early trial,
    such as a first run on a device still warming up, must not decide), and kept only when it
    beats the better of the start's two new scores by `min_gain`; the returned score is the
    kept schedule's new one. Returns (best cfg, its score, [(cfg, score), ...] in the order
    measured)."""
    space = SPACE if space is None else space
    best = dict(DEFAULT if start is None else start)
    trials, seen = [], {}

    def score(cfg):
        k = _key(cfg)
        if k not in seen:
            seen[k] = measure(dict(cfg))
            trials.append((dict(cfg), seen[k]))
            if log:
                log(f"tune: {cfg} -> {seen[k]}")
        return seen[k]

    best_score = score(best)
    if best_score is None:
        raise RuntimeError(f"the starting schedule {best} was rejected")
    for _ in range(passes):
        changed = False
        for field, values in space.items():
            for v in values:
                if v == best[field]:
                    continue
                cand = dict(best, **{field: v})
                s = score(cand)
                if s is not None and s > best_score * (1 + min_gain):
                    best, best_score, changed = cand, s, True
        if not changed:
            break
    start_cfg = dict(DEFAULT if start is None else start)
    if confirm and _key(best) != _key(start_cfg):
        again = [measure(dict(start_cfg)), measure(dict(best)), measure(dict(start_cfg))]
        trials += [(dict(start_cfg), again[0]), (dict(best), again[1]), (dict(start_cfg), again[2])]
        if log:
            log(f"tune: confirm start {again[0]}, {again[2]} against {best} {again[1]}")
        s0 = max(x for x in (again[0], again[2]) if x is not None) if again[0] or again[2] else None
        if again[1] is None or (s0 is not None and again[1] <= s0 * (1 + min_gain)):
            return start_cfg, s0, trials
        best_score = again[1]
    return best, best_score, trials


def tuning_of(cfg):
    """the gi_config_tuned fields of a schedule"""
    return {k: cfg[k] for k in TUNED_FIELDS if k in cfg}


def measure_batch(jobs, expected, reps=3, device=0):
    """A measure() for coordinate_search: decodes `jobs` (a ctypes array of gi_job, inputs and
    output buffers in device memory) with one gi_batch per candidate, once to warm up and then
    `reps` times, and returns the median GB/s of decoded bytes; None when any job's status,
    size, in_used or CRC-32 differs from expected[k] = (out_size, in_used, crc32), or when the
    candidate's contexts do not fit in free device memory."""
    from gpuzip_amd import gpuinflate as gi
    total = sum(e[0] for e in expected)

    def ok():
        for j, (n, used, crc) in zip(jobs, expected):
            if j.status != gi.GI_OK or j.out_size != n or j.in_used != used or j.crc32 != crc:
                return False
        return True

    def measure(cfg):
        try:
            b = gi.Batch(workers=cfg["workers"], device=device, segment_out_mb=cfg["segment_out_mb"],
                         tuning=tuning_of(cfg))
        except gi.GpuError:  # its buffers do not fit beside the caller's: not a candidate here
            return None
        with b:
            times = []
            for r in range(reps + 1):
                for j in jobs:
                    j.status = -1
                t0 = time.perf_counter()
                b.run(jobs)
                dt = time.perf_counter() - t0
                if not ok():
                    return None
                if r:
                    times.append(dt)
        return total / statistics.median(times) / 1e9

    return measure


def load(path, key):
    """the cached choice for `key` in the JSON file at `path`, or None"""
    try:
        with open(path) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def save(path, key, entry):
    """stores `entry` under `key` in the JSON file at `path` (written whole, then renamed)"""
    data = {}
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        pass
    data[key] = entry
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune_batch(jobs, expected, device=0, cache=None, key=None, space=None, passes=3, reps=3, log=None):
    """Chooses the batch schedule for `jobs` by measurement (coordinate_search over
    measure_batch), or takes it from the JSON file `cache` under `key` when it is there. Returns
    {"params": cfg, "GBs": score, "trials": n, "seconds": time spent, "cached": bool}."""
    if cache and key:
        hit = load(cache, key)
        if hit:
            return dict(hit, cached=True)
    t0 = time.perf_counter()
    best, best_score, trials = coordinate_search(measure_batch(jobs, expected, reps, device), space=space,
                                                 passes=passes, log=log)
    entry = {"params": best, "GBs": round(best_score, 3), "trials": len(trials),
             "rejected": sum(1 for _, s in trials if s is None), "seconds": round(time.perf_counter() - t0, 2),
             # the start's best score (it is measured again when a winner is confirmed)
             "default_GBs": round(max(x for c, x in trials if _key(c) == _key(trials[0][0]) and x is not None), 3)}
    if cache and key:
        save(cache, key, entry)
    return dict(entry, cached=False)
