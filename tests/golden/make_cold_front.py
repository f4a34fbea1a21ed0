"""Writes cold_front.json: the frozen cold-front matrix (25 x 25, stored float
coefficients) and right-hand side of the reference's
tests/test_gmres_precision.cpp:19-55, the only real-data matrix among the
reference's tests.  Numbers only; run in the build container, where the
reference is mounted:  python tests/golden/make_cold_front.py"""
import json
import pathlib
import re

SRC = pathlib.Path("/root/reference/tests/test_gmres_precision.cpp")
OUT = pathlib.Path(__file__).resolve().parent / "cold_front.json"


def main():
    text = SRC.read_text()
    body = text[text.index("entries[] = {"):]
    body = body[:body.index("};")]
    entries = [(int(r), int(c), v) for r, c, v in re.findall(r"\{(\d+),(\d+),([-+.\deE]+)f\}", body)]
    data = {
        "source": "tests/test_gmres_precision.cpp:19-55 (frozen CP2026 cold-front matrix)",
        "n": 25,
        "entries": [[r, c, v] for r, c, v in entries],  # values as written (float literals)
        "rhs_ones_at": list(range(0, 25, 5)),
        "restart": 25,
        "maximum_iterations": 200,
        "absolute_tolerance_in_float_eps": 32,
    }
    OUT.write_text(json.dumps(data, indent=1) + "\n")
    print(f"{OUT}: {len(entries)} entries")


if __name__ == "__main__":
    main()
