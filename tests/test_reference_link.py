"""The drop-in boundary at link time: the reference's own main.c, compiled
unchanged with its real main(), links against libgraphbench.so together with
the reference objects off the hot path, with no undefined symbol, and every
hot-path function main() calls (main/main.c:116-238) binds to the library.

Built by oracle/ref/Makefile (target `link`) from the reference sources; needs
the reference tree and MPICH, so it runs in the build container only.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
EXE = os.path.join(ROOT, "oracle", "_ref", "link", "ref_main")
LIB = os.path.join(ROOT, "graphbench_amd", "libgraphbench.so")

# the hot-path functions main.c calls, each replaced by the library
FROM_LIBRARY = ["run_sspath_arr", "run_pagerank_csr_mpi", "run_bc", "scan_csr_idx",
                "read_graph_csr", "comp_distrib_csr", "write_recip", "chunk_files",
                "remove_duplicate"]
# globals main.c defines (graph.h:79-91, main.c:8-22); the library's are weak
GLOBALS = ["num_threads", "num_nodes", "buck_size", "g_v_num", "bin", "buff_size",
           "prefetch_dis", "use_partition", "data_extend", "par_extend", "nodes_extension",
           "base_dir"]


@pytest.mark.skipif(not os.path.isdir(REF) or not os.path.exists("/opt/conda/include/mpi.h"),
                    reason="needs the reference sources and MPICH (build container only)")
def test_reference_main_links_against_library():
    if not os.path.exists(LIB):
        pytest.skip("libgraphbench.so not built")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref"), "link"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    dyn = subprocess.run(["nm", "-D", EXE], capture_output=True, text=True, check=True).stdout
    undef = {l.split()[-1] for l in dyn.splitlines() if l.split()[0] == "U"}
    for f in FROM_LIBRARY:
        assert f in undef, f"{f} does not bind to the shared library"
    lib_syms = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True,
                              text=True, check=True).stdout
    lib_defined = {l.split()[-1] for l in lib_syms.splitlines()}
    for f in FROM_LIBRARY:
        assert f in lib_defined, f
    # main.c's own definitions of the globals win over the library's weak ones
    full = subprocess.run(["nm", EXE], capture_output=True, text=True, check=True).stdout
    defined = {l.split()[-1] for l in full.splitlines() if len(l.split()) == 3
               and l.split()[1] in "BDbdGgSsC"}
    for gname in GLOBALS:
        assert gname in defined, f"global {gname} not defined by main.o"
    weak = {l.split()[-1] for l in lib_syms.splitlines() if l.split()[1] in "VvWw"}
    for gname in GLOBALS:
        assert gname in weak, f"library global {gname} is not weak"
