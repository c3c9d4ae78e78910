"""Runs the batched form's host logic on the CPU stand-in of the HIP
runtime (hipstub.cpp): every configuration of tests/test_gpu_batch.py in
file order, then the autotune sweep and a graph-captured run of its choice.
The stand-in's kernels compute nothing (film values are not checked); it
checks memory ranges of every launch, stream-capture rules (joins before
the end of a capture, no event from outside a capture, no memset node, no
synchronous call during a capture) and, under AddressSanitizer, every host
and "device" access.  Prints the stand-in's error count last."""
import ctypes as C
import pathlib
import sys

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))
import __graft_entry__ as ge  # noqa: E402

pkg = ge.load_package()
pkg.LIB_PATH = pathlib.Path(sys.argv[1])
wl = ge.load_workloads()
import test_gpu_batch as T  # noqa: E402

stub = C.CDLL(str(pkg.LIB_PATH))
for tuning in [dict(blocks_per_cu=0, tile_waves=1, use_graph=True), dict(blocks_per_cu=1, tile_waves=4, use_graph=False),
               dict(blocks_per_cu=2, tile_waves=2, use_graph=True), dict(streams=1, use_graph=False),
               dict(streams=3, use_graph=True), dict(streams=4, use_graph=True), dict(streams=8, use_graph=True)]:
    try:
        T.test_batch_equals_plain(pkg, wl, tuning)
    except AssertionError:
        pass
try:
    T.test_batch_mixed_lights_and_zero_spp(pkg, wl)
except AssertionError:
    pass
for n_jobs in (6, 2):
    _, jobs, scenes = T._small_batch(pkg, wl, n_jobs=n_jobs)
    b = T._batch(pkg, jobs, scenes)
    best = b.autotune(reps=2)
    print("autotune", n_jobs, best, flush=True)
    b.run()
    for j in range(len(jobs)):
        b.film(j)
print("STUB_ERRORS", stub.stub_errors(), flush=True)
