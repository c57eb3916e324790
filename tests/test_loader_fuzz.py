"""Seeded mutation fuzz of the host readers (t/v/e text graphs, the binary CSR cache, .card maps)
under AddressSanitizer and UndefinedBehaviorSanitizer.  The readers are built with g++ from their
own sources, without the HIP units, so this runs on any host.  Every mutated file must come back
with a status code and no sanitizer report; valid seeds must still load."""
import os
import random
import shutil
import struct
import subprocess

import pytest

from tests import common as T

ROOT = os.path.join(os.path.dirname(__file__), "..")
CSRC = os.path.join(ROOT, "asdmatch_amd", "csrc")
DRIVER = os.path.join(os.path.dirname(__file__), "fuzz", "loader_fuzz.cpp")
CARDS = os.path.join(T.GOLDEN, "yeast", "cards")
QUERIES = os.path.join(T.GOLDEN, "yeast", "queries")


@pytest.fixture(scope="module")
def fuzz_bin(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("fuzz") / "loader_fuzz")
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, DRIVER, os.path.join(CSRC, "graph_host.cpp"),
           os.path.join(CSRC, "card_io.cpp"), "-o", out, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    if r.returncode != 0 and "asan" in r.stderr.lower():
        pytest.skip("sanitizer runtime not available: " + r.stderr[-300:])
    assert r.returncode == 0, r.stderr[-2000:]
    return out


def _env():
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:verify_asan_link_order=0:abort_on_error=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1"
    return env


def _run(binary, mode, files):
    r = subprocess.run([binary, mode] + files, capture_output=True, text=True, timeout=600, env=_env())
    assert r.returncode == 0 and "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]
    return {int(k): int(v) for k, v in (ln[3:].split(":") for ln in r.stdout.split("\n") if ln.startswith("rc "))}


NUMBERS = [0, 1, 2, 5, 255, 1 << 16, (1 << 24) - 1, 1 << 24, 1 << 31, (1 << 32) - 1, 1 << 32, 10 ** 12]


def _mutate_text(R, data):
    lines = data.split(b"\n")
    for _ in range(R.randint(1, 4)):
        if not lines:
            lines = [b"t 1 0"]
        i = R.randrange(len(lines))
        k = R.randrange(8)
        if k == 0:
            del lines[i]
        elif k == 1:
            lines.insert(i, lines[R.randrange(len(lines))])
        elif k == 2:  # replace one number by an extreme value
            w = lines[i].split(b" ")
            w[R.randrange(len(w))] = str(R.choice(NUMBERS + [R.randrange(100)])).encode()
            lines[i] = b" ".join(w)
        elif k == 3:
            lines[i] = lines[i][:R.randrange(len(lines[i]) + 1)]
        elif k == 4:
            j = R.randrange(len(lines))
            lines[i], lines[j] = lines[j], lines[i]
        elif k == 5:
            lines[i] = bytes(R.choice(b"vet0123456789 \t-x") for _ in range(R.randint(0, 12)))
        elif k == 6:
            lines = lines[:i]
        else:
            lines.insert(i, R.choice([b"e 0 0", b"v 0 0 0", b"e 1 1", b"t 3 1", b"v 99999 1 1"]))
    return b"\n".join(lines)


def _mutate_bytes(R, data, keep=0):
    b = bytearray(data)
    for _ in range(R.randint(1, 3)):
        k = R.randrange(4)
        if k == 0 and len(b) > keep:
            b[R.randrange(keep, len(b))] = R.randrange(256)
        elif k == 1 and len(b) >= keep + 4:
            i = R.randrange(keep, len(b) - 3)
            b[i:i + 4] = struct.pack("<I", R.choice([0, 1, 1 << 24, (1 << 32) - 1, R.randrange(50)]))
        elif k == 2:
            b = b[:R.randrange(len(b) + 1)]
        else:
            b += bytes(R.randrange(256) for _ in range(R.randint(1, 16)))
    return bytes(b)


def test_seeds_load_cleanly(fuzz_bin):
    seeds = sorted(os.path.join(QUERIES, f) for f in os.listdir(QUERIES))[:20]
    assert _run(fuzz_bin, "g", seeds) == {0: len(seeds)}
    cards = sorted(os.path.join(CARDS, f) for f in os.listdir(CARDS) if f.endswith(".card"))
    assert _run(fuzz_bin, "c", cards) == {0: len(cards)}


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_mutated_graphs_and_cards_fail_cleanly(fuzz_bin, tmp_path, seed):
    R = random.Random(seed)
    seeds = sorted(os.path.join(QUERIES, f) for f in os.listdir(QUERIES))[:40]
    seeds += sorted(os.path.join(CARDS, f) for f in os.listdir(CARDS) if f.endswith(".graph"))
    texts = [open(p, "rb").read() for p in seeds]
    bins = []
    for i, p in enumerate(seeds[:8]):
        out = str(tmp_path / f"seed{i}.csr")
        assert subprocess.run([fuzz_bin, "s", out, p], env=_env(), timeout=60).returncode == 0
        bins.append(open(out, "rb").read())
    files = []
    for t in range(300):
        p = str(tmp_path / f"g{t}")
        if R.random() < 0.3:
            data = _mutate_bytes(R, R.choice(bins), keep=8)  # keep the binary header's magic
        else:
            data = _mutate_text(R, R.choice(texts))
        open(p, "wb").write(data)
        files.append(p)
    codes = _run(fuzz_bin, "g", files)
    assert sum(codes.values()) == len(files)
    assert set(codes) <= {0, -2, -7}, codes  # loaded, malformed, or beyond the supported label range
    cards = [open(os.path.join(CARDS, f), "rb").read() for f in sorted(os.listdir(CARDS)) if f.endswith(".card")]
    cfiles = []
    for t in range(300):
        p = str(tmp_path / f"c{t}")
        open(p, "wb").write(_mutate_bytes(R, R.choice(cards)))
        cfiles.append(p)
    codes = _run(fuzz_bin, "c", cfiles)
    assert sum(codes.values()) == len(cfiles)
    assert set(codes) <= {0, -1, -2}, codes
