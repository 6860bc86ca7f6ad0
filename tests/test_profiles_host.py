"""The committed rocprofv3 summaries under profiles/ and bench.py's reading of
them (CPU only: no GPU, no profiler).

bench.py's roofline takes the per-launch HBM bytes of the J/K kernels from
profiles/eri_jk64_hbm_traffic_592bf.txt.  These tests check that the file's
machine-readable lines follow from its own counter table and copy calibration,
that bench.py's parser returns their sum, and that the kernel-trace summary
names the dominant kernel whose average it quotes beside bench.py's HIP-event
figure.  They also check that every summary carries the full rocprofv3
command that made it, and that bench.py refuses more than one GPU before it
touches a device."""
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROFILES = os.path.join(ROOT, "profiles")
TRAFFIC = os.path.join(PROFILES, "eri_jk64_hbm_traffic_592bf.txt")
TRACE = os.path.join(PROFILES, "eri_fock_build_592bf_kernel_trace.txt")


def _bench():
    sys.path.insert(0, ROOT)
    import bench
    return bench


def _symbol(s):
    return s.replace(" ", "")


def _traffic_table():
    """(scale_fetch, scale_write, rows {symbol: (fetch KiB, write KiB, bytes)},
    traffic {class: bytes}, kernel {class: symbol}) of the PMC summary."""
    scale_f = scale_w = None
    rows, traffic, kernel = {}, {}, {}
    row_re = re.compile(r"^#\s+(eri_\w+<[^>]*>)\s+(\d+)\s+(\d+)\s+([0-9.e+]+)\s*$")
    with open(TRAFFIC) as f:
        for line in f:
            m = re.search(r"FETCH_SIZE .*-> scale ([0-9.]+);\s+WRITE_SIZE .*-> scale ([0-9.]+)", line)
            if m:
                scale_f, scale_w = float(m.group(1)), float(m.group(2))
                continue
            m = row_re.match(line)
            if m:
                rows[_symbol(m.group(1))] = (int(m.group(2)), int(m.group(3)), float(m.group(4)))
                continue
            w = line.split()
            if len(w) == 3 and w[0] == "traffic":
                traffic[w[1]] = float(w[2])
            elif len(w) == 3 and w[0] == "kernel":
                kernel[w[1]] = w[2]
    return scale_f, scale_w, rows, traffic, kernel


def test_traffic_lines_follow_from_the_counter_table():
    scale_f, scale_w, rows, traffic, kernel = _traffic_table()
    # the copy kernel reads and writes 1 GiB per launch; FETCH_SIZE counts half of the read
    assert scale_f == pytest.approx(2.0, rel=1e-3) and scale_w == pytest.approx(1.0, rel=1e-3)
    assert len(rows) == len(traffic) == len(kernel) >= 20
    assert set(traffic) == set(kernel)
    for cls, sym in kernel.items():
        assert re.fullmatch(r"[0-2]{4}", cls)
        # the symbol's template arguments are the class's angular momenta
        assert sym.split("<")[1].split(",")[:4] == list(cls), (cls, sym)
        fetch, write, table_bytes = rows[sym]
        recomputed = (fetch * scale_f + write * scale_w) * 1024
        assert recomputed == pytest.approx(table_bytes, rel=2e-3), sym
        assert traffic[cls] == pytest.approx(recomputed, rel=1e-4), sym


def test_bench_reads_the_build_traffic_it_reports():
    bench = _bench()
    _, _, _, traffic, kernel = _traffic_table()
    total = bench.build_traffic_from_profile()
    assert total == pytest.approx(sum(traffic.values()), rel=1e-12)
    assert 1e10 < total < 1e11          # tens of GB per build: the J/K accumulation dominates
    for cls in traffic:
        assert bench.traffic_from_profile(cls) == traffic[cls]
        assert bench.kernel_from_profile(cls) == kernel[cls]
    assert bench.traffic_from_profile("3333") is None


def _trace_rows():
    rows = {}
    row_re = re.compile(r"^\s+(\w+<[^>]*>)\s+(\d+)\s+([0-9.]+)\s+([0-9.]+)\s+([0-9.]+)\s+([0-9.]+)\s+([0-9.]+)\s*$")
    quoted = None
    with open(TRACE) as f:
        for line in f:
            m = row_re.match(line)
            if m:
                calls, total, avg, lo, hi = (int(m.group(2)),) + tuple(float(m.group(i)) for i in range(3, 7))
                rows[_symbol(m.group(1))] = (calls, total, avg, lo, hi)
                continue
            m = re.search(r"for (\S+): kernel_ms ([0-9.]+) \(rocprof avg ([0-9.]+)", line)
            if m:
                quoted = (m.group(1), float(m.group(2)), float(m.group(3)))
    return rows, quoted


def test_kernel_trace_names_the_dominant_kernel_it_quotes():
    rows, quoted = _trace_rows()
    assert quoted is not None
    sym, kernel_ms, rocprof_ms = quoted
    # the J/K kernels of the builds (the Schwarz-bound kernels run once, at set-up)
    eri = {k: v for k, v in rows.items() if k.startswith("eri_") and "schwarz" not in k}
    assert len(eri) >= 16
    for k, (calls, total, avg, lo, hi) in eri.items():
        assert calls == 12                              # 2 warmup + 10 timed concurrent builds
        assert avg == pytest.approx(total / calls, rel=1e-3), k
        assert lo <= avg <= hi, k
    dominant = max(eri, key=lambda k: eri[k][2])
    assert dominant == sym
    assert eri[sym][2] / 1e3 == pytest.approx(rocprof_ms, abs=1e-3)
    # bench.py's HIP-event average over the timed builds agrees with the profiler's
    assert abs(kernel_ms - rocprof_ms) / rocprof_ms < 0.05


def test_every_summary_names_its_command():
    names = sorted(n for n in os.listdir(PROFILES) if not n.startswith("."))
    assert names
    for n in names:
        assert n.endswith((".txt", ".md")), n
        with open(os.path.join(PROFILES, n)) as f:
            text = f.read()
        assert "rocprofv3 " in text and " -- python3 " in text, n
        assert "_work/" not in text, n
        assert len(text) < 16384, n


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
