"""CPU build of the device geometry checked bit-for-bit against the oracle.

d2_geom.hpp is the arithmetic of every kernel; compiling it for the host and
comparing with the oracle on random inputs pins the kernels' FP64 expression
order without a GPU (tri-tri distance, point-triangle CCD with the per-pair
frame changes, edge-edge CCD incl. bisection cases, state extrapolation), and
the subtree-pruning bound never hides a leaf contact, and the FP32 culls
(CCD soup sub-tests, discrete triangle pairs at the hit threshold) never
drop a sub-test or pair that reports a hit in FP64."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_device_geometry_matches_oracle_bitwise(tmp_path):
    exe = tmp_path / "cpu_geom_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", os.path.join(ROOT, "tests", "cpu_geom_check.cpp"),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    rows = {l.split()[0]: l.split() for l in out.strip().splitlines()}
    for name in ("tri_tri_dist", "point_tri", "edge_edge", "extrapolate", "near_mid"):
        assert int(rows[name][2]) == 0, out
        assert int(rows[name][1]) >= 10000
    assert "bisection cases" in out
    # subtree pruning: no violation, and the random placements do prune often
    assert int(rows["nodes_apart"][2]) == 0, out
    line = next(l for l in out.splitlines() if l.startswith("nodes_apart"))
    prunes = [int(x) for x in re.findall(r"(?:boxes|spheres|f32|axes) (\d+)", line)]  # boxes, spheres, f32, face axes
    assert len(prunes) == 4 and min(prunes) > 1000, out
    assert prunes[3] <= prunes[0], out  # six of nodes_apart's fifteen axes
    # FP32 culls of the CCD soup sub-tests: no culled sub-test reports a hit,
    # on draws with thousands of hits and most sub-tests culled
    assert int(rows["soup_culls"][2]) == 0, out
    line = next(l for l in out.splitlines() if l.startswith("soup_culls"))
    culled, hits = (int(x) for x in re.findall(r"(?:culled|hits) (\d+)", line))
    assert int(rows["soup_culls"][1]) >= 100000 and culled > int(rows["soup_culls"][1]) // 2 and hits > 1000, out
    # the triangle-level edge cull (nine edge-edge sub-tests at once) is sound
    assert int(rows["tri_edges"][2]) == 0, out
    assert int(re.search(r"tri_edges .*culled (\d+)", out).group(1)) > 1000, out
    # the point-triangle cull's lateral bound holds the reference's float
    # barycentrics: points just outside an edge of thin triangles
    assert int(rows["edge_points"][2]) == 0, out
    culled, hits = (int(x) for x in re.search(r"edge_points .*culled (\d+), hits (\d+)", out).groups())
    assert culled > 1000 and hits > 100, out
    # FP32 classification of discrete triangle pairs near the hit threshold:
    # no certain miss hits, every certain hit hits; each outcome occurs often
    assert int(rows["disc_culls"][2]) == 0, out
    line = next(l for l in out.splitlines() if l.startswith("disc_culls"))
    counts = [int(x) for x in re.findall(r"(?:miss|hit|fp64 hits) (\d+)", line)]
    assert int(rows["disc_culls"][1]) >= 20000 and len(counts) == 3 and min(counts) > 1000, out


def test_partition_rank_sort_matches_std_sort(tmp_path):
    """The post-processing kernel replays std::sort (the contact filter's
    unstable sort, contact.h:44-69) by partition ranks; the formulation is
    checked against libstdc++ on tie-heavy keys."""
    exe = tmp_path / "cpu_sort_check"
    subprocess.run(["g++", "-O2", "-std=c++17", os.path.join(ROOT, "tests", "cpu_sort_check.cpp"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert out[0] == "sort" and int(out[1]) >= 1000 and int(out[2]) == 0, out
