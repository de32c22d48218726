"""Writes tests/golden/rescue_golden.json: for every case of tests/rescue_cases.py, what the
reference's own rescue (gpuinflate/rescue.cpp, built from the reference sources into
oracle/_ref/libgrref.so by oracle/Makefile) returns - the code, SHA-256 of the output and of the
mask, the statistics, the pieces and the notes. The parity tests check against these where the
reference library is absent.

    python tests/golden/make_rescue_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import rescue_cases  # noqa: E402
import rescue_lib  # noqa: E402


def main():
    ref = rescue_lib.reference()
    if ref is None:
        sys.exit("oracle/_ref/libgrref.so is missing: run make -C oracle where the reference sources are")
    golden = {}
    for name, data, kw in rescue_cases.small_cases():
        golden[name] = rescue_lib.summary(rescue_lib.rescue(ref, data, **kw))
        print(name, golden[name]["rc"], golden[name]["out_len"], flush=True)
    with open(os.path.join(HERE, "rescue_golden.json"), "w") as f:
        json.dump(golden, f, indent=0, sort_keys=True)


if __name__ == "__main__":
    main()
