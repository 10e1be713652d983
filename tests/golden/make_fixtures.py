"""Writes the golden fixtures in this folder.

Reference-held data (read as text from the reference tree, which must be present):
  ranlxs0_seed32767.json  -- Luscher's ranlxs test vector from tests/ranlx/main.C:40-135
  ludcmp_known_answer.json -- matrix, det and log det from tests/ludcmp_test/main.C:18-65
Published data not in the reference:
  philox4x32_10_kat.json   -- Random123 known-answer vectors (Salmon et al., SC'11)
"""
import json
import os
import re

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def ranlxs():
    src = open(os.path.join(REF, "tests/ranlx/main.C")).read()
    vals = {int(k): float(v) for k, v in re.findall(r"xsn\[(\d+)\]=([0-9.]+)f;", src)}
    assert sorted(vals) == list(range(96))
    return {
        "source": "tests/ranlx/main.C:28-37,40-135,137-150",
        "level": 0, "seed": 32767, "batches": 20, "batch_size": 204, "offset_in_last_batch": 60,
        "scale": 2 ** 24,
        "values": [vals[k] for k in range(96)],
    }


def ludcmp():
    src = open(os.path.join(REF, "tests/ludcmp_test/main.C")).read()
    det = re.search(r"det = ([0-9.]+) \* 10\^(-?\d+)", src)
    ldet = re.search(r"ldet = (-?[0-9.]+)", src)
    return {
        "source": "tests/ludcmp_test/main.C:18-33,52,65",
        "n": 10, "matrix": "1/(s+t+2)",
        "det": float(det.group(1)) * 10.0 ** int(det.group(2)),
        "det_text": det.group(0), "logdet": float(ldet.group(1)),
    }


def philox():
    return {
        "source": "Random123 kat_vectors, philox4x32 10 rounds",
        "vectors": [
            {"ctr": [0, 0, 0, 0], "key": [0, 0],
             "out": [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]},
            {"ctr": [0xffffffff] * 4, "key": [0xffffffff] * 2,
             "out": [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]},
            {"ctr": [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
             "key": [0xa4093822, 0x299f31d0],
             "out": [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]},
        ],
    }


if __name__ == "__main__":
    for name, fn in [("ranlxs0_seed32767.json", ranlxs), ("ludcmp_known_answer.json", ludcmp),
                     ("philox4x32_10_kat.json", philox)]:
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(fn(), f, indent=1)
        print("wrote", name)
