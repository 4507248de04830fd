"""Regenerates the golden scan fixtures from the reference's shipped data files.

Inputs (data, not source) read from the reference checkout, only in the build
container:
  samples/grep_text/input/pg100.txt.small   20,888 bytes of text
  samples/grep_text/input/dict.gpu          58,112 NUL-padded 32-byte words
Outputs in this directory:
  pg100_small.txt      the text, copied as-is
  dict_subset.bin      every dictionary word that occurs in the text as a token
                       (case-sensitive), plus the dictionary's duplicated words
                       and its one word holding a non-word byte, plus a few
                       hand-added edge records (empty, 32-character, uppercase)
  pg100_counts.json    per-word counts for window sizes 16384 (the reference's
                       CORPUS_PREFETCH_SIZE) and 64, from the literal restatement
                       of match_string (oracle.orc_scan_naive)
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference/samples/grep_text/input"
sys.path.insert(0, ROOT)
from oracle import oracle as orc  # noqa: E402


def main():
    text = open(os.path.join(REF, "pg100.txt.small"), "rb").read()
    d = open(os.path.join(REF, "dict.gpu"), "rb").read()
    recs = [d[i * 32:(i + 1) * 32] for i in range(len(d) // 32)]
    words = [r.split(b"\0")[0] for r in recs]
    tokens = set(re.findall(rb"[A-Za-z_]+", text))
    seen, dups = set(), set()
    for w in words:
        if w in seen:
            dups.add(w)
        seen.add(w)
    keep = []
    for r, w in zip(recs, words):
        if w in tokens or w in dups or not re.fullmatch(rb"[A-Za-z_]*", w):
            keep.append(r)
    # duplicates of kept words stay duplicated (the dictionary has 2 such pairs)
    extra = [b"", b"The", b"A", b"a" * 32, b"Project", b"Gutenberg", b"the", b"the"]
    for w in extra:
        keep.append(w.ljust(32, b"\0"))
    keep = keep[:1024]
    dict_arr = np.frombuffer(b"".join(keep), dtype=np.uint8).reshape(-1, 32)
    out = {"n_words": len(keep), "counts": {}}
    for chunk in (16384, 64):
        out["counts"][str(chunk)] = [int(x) for x in orc.scan_naive(text, dict_arr, chunk)]
    open(os.path.join(HERE, "pg100_small.txt"), "wb").write(text)
    open(os.path.join(HERE, "dict_subset.bin"), "wb").write(dict_arr.tobytes())
    json.dump(out, open(os.path.join(HERE, "pg100_counts.json"), "w"))
    print("words", len(keep), "total matches", {k: sum(v) for k, v in out["counts"].items()})


if __name__ == "__main__":
    main()
