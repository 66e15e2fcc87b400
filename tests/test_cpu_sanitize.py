"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer:
the Matrix Market reader/writer (hala_amd/csrc/mmio.cpp) on valid files of
every format, field and symmetry and on a few hundred seeded mutations of them
(truncated, reordered and corrupted lines, huge, negative and out-of-range
sizes and indices, non-numeric tokens, overlong lines), and the synthetic
generators (hala_amd/csrc/stencils.cpp) on small and degenerate sizes.  Built
with g++ for the host only (tests/cpp/host_sanitize.cpp); any sanitizer report
aborts the driver, and every input must end in a status, not a crash."""
import os
import random
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SEEDS = {
    "coord_real_general": "%%MatrixMarket matrix coordinate real general\n% c\n4 5 6\n1 1 4.0\n2 3 -1.5\n3 2 2e-3\n4 5 1\n1 5 -7\n4 1 0.5\n",
    "coord_int_sym": "%%MatrixMarket matrix coordinate integer symmetric\n3 3 4\n1 1 2\n2 1 -1\n3 2 -1\n3 3 2\n",
    "coord_pattern_general": "%%MatrixMarket matrix coordinate pattern general\n3 4 3\n1 2\n2 4\n3 1\n",
    "coord_real_skew": "%%MatrixMarket matrix coordinate real skew-symmetric\n4 4 3\n2 1 1.0\n3 1 -2.0\n4 3 3.5\n",
    "array_real_general": "%%MatrixMarket matrix array real general\n2 3\n1\n2\n3\n4\n5\n6\n",
    "array_real_sym": "%%MatrixMarket matrix array real symmetric\n3 3\n1\n2\n3\n4\n5\n6\n",
    "array_int_skew": "%%MatrixMarket matrix array integer skew-symmetric\n3 3\n1\n2\n3\n",
}

TOKENS = ["", "-1", "0", "2147483647", "2147483648", "99999999999999999999", "-99999999999999999999", "nan",
          "inf", "1e400", "x", "%", "3.5", "1 1", "\t", "+"]


def mutate(text, rng):
    lines = text.split("\n")
    kind = rng.randrange(9)
    if kind == 0 and len(lines) > 2:  # truncate
        lines = lines[: rng.randrange(1, len(lines))]
    elif kind == 1 and len(lines) > 3:  # drop a line
        del lines[rng.randrange(1, len(lines) - 1)]
    elif kind == 2:  # duplicate a line
        k = rng.randrange(len(lines))
        lines.insert(k, lines[k])
    elif kind == 3:  # replace a token
        k = rng.randrange(len(lines))
        toks = lines[k].split(" ")
        toks[rng.randrange(len(toks))] = rng.choice(TOKENS)
        lines[k] = " ".join(toks)
    elif kind == 4:  # overlong line
        lines.insert(rng.randrange(1, len(lines)), "1 " * rng.choice([2100, 5000]))
    elif kind == 5:  # corrupt bytes
        b = bytearray("\n".join(lines).encode())
        for _ in range(rng.randrange(1, 6)):
            b[rng.randrange(len(b))] = rng.randrange(1, 256)
        return bytes(b)
    elif kind == 6:  # size line: random sizes
        for k, l in enumerate(lines):
            if k > 0 and l and not l.startswith("%"):
                n = len(l.split())
                lines[k] = " ".join(rng.choice(TOKENS[1:7] + ["3", "5", "6"]) for _ in range(n))
                break
    elif kind == 7:  # banner words
        toks = lines[0].split(" ")
        toks[rng.randrange(len(toks))] = rng.choice(["complex", "hermitian", "vector", "MATRIX", "Coordinate", ""])
        lines[0] = " ".join(toks)
    else:  # swap two entry lines
        if len(lines) > 4:
            a, b = rng.randrange(2, len(lines)), rng.randrange(2, len(lines))
            lines[a], lines[b] = lines[b], lines[a]
    return "\n".join(lines).encode()


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    out = tmp_path_factory.mktemp("san") / "host_sanitize"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", f"-I{ROOT}/include", f"{ROOT}/hala_amd/csrc/mmio.cpp",
           f"{ROOT}/hala_amd/csrc/stencils.cpp", f"{ROOT}/tests/cpp/host_sanitize.cpp", "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "sanitizer" in r.stderr.lower() and "cannot find" in r.stderr.lower():
        pytest.skip("sanitizer runtime not installed")
    assert r.returncode == 0, r.stderr[-3000:]
    return out


def test_host_code_under_sanitizers(driver, tmp_path):
    rng = random.Random(20261203)
    files = []
    for name, text in SEEDS.items():
        f = tmp_path / f"{name}.mtx"
        f.write_text(text)
        files.append(f)
        for m in range(40):
            g = tmp_path / f"{name}_m{m}.mtx"
            g.write_bytes(mutate(text, rng))
            files.append(g)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(driver), str(tmp_path / "roundtrip.mtx")] + [str(f) for f in files], capture_output=True,
                       text=True, env=env, timeout=300)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    status = dict(l.rsplit(" ", 1) for l in r.stdout.splitlines() if l)
    assert status.pop("generators") == "ok"
    assert len(status) == len(files)
    # the valid seeds read, convert and round-trip; some mutations are refused
    for name in SEEDS:
        assert status[str(tmp_path / f"{name}.mtx")] == "0", name
    assert sum(v != "0" for v in status.values()) > 50
