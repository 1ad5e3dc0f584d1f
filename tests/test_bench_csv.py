"""The reference-schema CSV writer of bench.py (host logic only, no GPU).

Schema and throughput rule: src/main.cu:300 and src/common_code.cuh:206-213
-- one line per timed launch, throughput = used_data_size / (t / 1000) / 2^30
with used_data_size counting 8 bytes per element per input array (plus
(n+7)/8 mask bytes for compaction, src/benchs/bench_compressstore_slabs.cu:108-111)."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_csv_schema_and_throughput(tmp_path):
    import bench
    n = 1 << 27
    ops = [
        {"op": "add", "w": 11, "in": "slab", "n": n, "times": [1.120, 1.130]},
        {"op": "sum", "w": 4, "in": "bitstream", "n": n, "times": [0.5]},
        {"op": "compressstore(mask)", "w": 11, "sel": 0.5, "mask": "uniform", "in": "slab", "n": n, "times": [0.901]},
        {"op": "compressstore(mask)", "w": 11, "sel": 0.01, "mask": "ref_1_cluster", "in": "slab", "n": n,
         "times": [0.251]},
        {"op": "groupby sum+count", "n": n, "key_bits": 10, "times": [1.276]},
        {"op": "matmul u7 8192^3 (unpack+gemm+pack)", "ms": 0.58},  # not a reference row
    ]
    files = bench.write_csv(ops, str(tmp_path))
    assert files == ["comp/sc_0.01", "comp/uni_0.5", "sweep"]
    lines = open(tmp_path / "sweep").read().splitlines()
    assert lines[0] == "approach;element_count;bits;block_count;thread_count;time_ms;throughput;run"
    f = lines[1].split(";")
    assert f[:5] == ["binary_op_slabs", str(n), "11", "-1", "-1"] and f[7] == "0"
    # the reference's published add row: 2 GiB in 1.120 ms -> 1785.7 GiB/s (data/1024_mb_rtx8000:39801)
    assert abs(float(f[6]) - 2 * n * 8 / (1.120e-3) / 2**30) < 1e-3 and abs(float(f[6]) - 1785.71) < 0.01
    assert lines[2].split(";")[7] == "1"
    assert lines[3].startswith("hop_nogaps;") and lines[4].startswith("groupby_sm_array;")
    c = open(tmp_path / "comp" / "uni_0.5").read().splitlines()[1].split(";")
    assert c[0] == "compressstore_slabs"
    assert abs(float(c[6]) - (8 * n + n // 8) / 0.901e-3 / 2**30) < 1e-3
