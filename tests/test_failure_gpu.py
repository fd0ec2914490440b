"""Failure paths, forced from the host: the negative-control build
(lib/libcft_plants.so, the same kernels with host code built -DCFT_PLANTS)
opened with CFT_PLANT=fail-launch refuses a handle's first compute launch
before anything is launched, as a failed launch would be reported.  Every
kernel that does run exits normally; nothing here faults the GPU.

Checked, in a child process (the plant library loads in place of
libcft.so): the call that meets the failure returns CFT_ERR_INTERNAL and
poisons the handle, every later call on it is refused, a fresh handle
without the plant computes the oracle's bits, an unknown plant name is
refused at open, and the serving loop fails exactly the job whose launch
failed - with a status, no hang - and serves the next one correctly."""

import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import PKG, oracle_elem

pytestmark = pytest.mark.gpu

PLANTS = os.path.join(PKG, "lib", "libcft_plants.so")

CHILD = r'''
import json, os, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from cftamd import CftError, Device, _lib, workload
from cftamd import serve as S
out = {}
n = 4096
a, b, c = workload.elementwise(2, n, seed=11)

os.environ["CFT_PLANT"] = "fail-launch"
dev = Device("hip:0")
try:
    dev.run(0, 3, 0, a, b, c)
    out["first"] = "ok"
except CftError as e:
    out["first"] = [e.status, str(e)]
try:
    dev.run(0, 3, 0, a, b, c)
    out["second"] = "ok"
except CftError as e:
    out["second"] = [e.status, str(e)]
dev.close()

os.environ["CFT_PLANT"] = "no-such-plant"
try:
    Device("hip:0").close()
    out["unknown"] = "ok"
except CftError as e:
    out["unknown"] = [e.status, str(e)]

os.environ["CFT_PLANT"] = ""
clean = Device("hip:0")
d, fl = clean.run(0, 3, 0, a, b, c)
out["clean"] = [d.tobytes().hex(), fl]

# the server's workers open their own handles: with the plant, the first
# launch of the elementwise worker fails
os.environ["CFT_PLANT"] = "fail-launch"
bufs = [clean.alloc(n * 32) for _ in range(5)]
for buf, x in zip(bufs, (a, b, c)):
    buf.host[:] = x
for buf in bufs:
    buf.to_device()
srv = S.Server(clean, workers=2)
jobs = []
for k in (3, 4):
    t = srv.submit_elem(_lib.JOB_ELEM, 0, 3, 0, bufs[0].data_ptr, bufs[1].data_ptr,
                        bufs[2].data_ptr, bufs[k].data_ptr, n)
    r = srv.wait(t)
    jobs.append([r.status, r.flags])
srv.close()
bufs[4].from_device()
out["served"] = jobs
out["served_second"] = bufs[4].host.tobytes().hex()
for buf in bufs:
    buf.free()
clean.close()
print("RESULT " + json.dumps(out), flush=True)
'''


def test_failed_launch_poisons_and_fails_the_job(oracle):
    assert os.path.exists(PLANTS), "lib/libcft_plants.so is not built"
    env = dict(os.environ, CFT_LIB=PLANTS)
    p = subprocess.run([sys.executable, "-c", CHILD, PKG], env=env, capture_output=True,
                       text=True, timeout=180)
    assert p.returncode == 0, p.stderr[-2000:]
    line = [x for x in p.stdout.splitlines() if x.startswith("RESULT ")][-1]
    out = json.loads(line[7:])
    internal, invalid = 8, 1
    assert out["first"][0] == internal and "poisoned" in out["first"][1], out["first"]
    assert out["second"][0] == internal and "GPU failure" in out["second"][1], out["second"]
    assert out["unknown"][0] == invalid and "no-such-plant" in out["unknown"][1]
    from cftamd import workload
    a, b, c = workload.elementwise(2, 4096, seed=11)
    want, wfl = oracle_elem(oracle, 3, 0, 0, a, b, c)
    assert bytes.fromhex(out["clean"][0]) == want.tobytes() and out["clean"][1] == wfl
    # the served job whose launch failed has a status; the next is served
    assert out["served"][0][0] == internal, out["served"]
    assert out["served"][1] == [0, wfl], out["served"]
    assert bytes.fromhex(out["served_second"]) == want.tobytes()
