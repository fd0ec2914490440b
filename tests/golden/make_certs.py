"""Copy the reference's ODE certificate data into tests/golden/certs/.

    python tests/golden/make_certs.py [REFERENCE_ROOT]

The reference certifies three ODE programs (Lorenz-63 RK4, Lorenz-96 RK4 at
ring 40, Henon-Heiles leapfrog) in binary64 and binary256 by the states at
every segment boundary of a run: certificates/<system>-<fmt>/ holds the
certificate text and states/run-R-boundary-K.bin, each the lanes' scratch
blocks (lanes x slots x element bytes).  This script copies, for runs 0
("main", the classic bank) and 1 ("half-step", the bank with its h slots
halved, run-1.bank), the boundary states and the run-1 bank, and writes
certs.json with what the certificate states about them: the image hash, the
digest of image and bank, lanes, steps a segment, and for each segment the
hash of its start and end states (SHA-256 of a domain tag and the state
bytes) and its flags and status.

Run 2 of the binary64 certificates ("wider") uses an image the reference
generates elsewhere and is not copied.  The files are data: inputs and
expected outputs.
"""

import hashlib
import json
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SYSTEMS = [("lorenz63-rk4", "fp256"), ("lorenz96-rk4", "fp256"), ("henonheiles-lf", "fp256"),
           ("lorenz63-rk4", "fp64"), ("lorenz96-rk4", "fp64"), ("henonheiles-lf", "fp64")]


# An open certificate's state hash: SHA-256 of the tag "cft-certificate 1
# state" with its terminating NUL, then the state bytes (host/tools/
# cert_write.h:80-90, 148-186).
STATE_TAG = b"cft-certificate 1 state\x00"


def state_hash(data):
    return hashlib.sha256(STATE_TAG + data).hexdigest()


def parse_cert(text):
    runs = []
    cur = None
    for line in text.splitlines():
        f = line.split()
        if not f:
            continue
        if f[0] == "run":
            cur = {"index": int(f[1]), "kind": f[2], "segments": []}
            runs.append(cur)
        elif cur is None:
            continue
        elif f[0] in ("program-image", "program-digest"):
            cur[f[0].replace("-", "_")] = f[1]
        elif f[0] in ("lanes", "steps"):
            cur[f[0]] = int(f[1])
        elif f[0] == "segment":
            cur["segments"].append({"start": f[3], "end": f[5], "flags": int(f[7]),
                                    "status": int(f[9])})
    return runs


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out_root = os.path.join(HERE, "certs")
    os.makedirs(out_root, exist_ok=True)
    index = {}
    for system, fmt in SYSTEMS:
        name = f"{system}-{fmt}"
        src = os.path.join(ref, "certificates", name)
        with open(os.path.join(src, name + ".cert")) as fh:
            runs = [r for r in parse_cert(fh.read()) if r["index"] in (0, 1)]
        dst = os.path.join(out_root, name)
        os.makedirs(dst, exist_ok=True)
        shutil.copyfile(os.path.join(src, "run-1.bank"), os.path.join(dst, "run-1.bank"))
        for r in runs:
            for k in range(len(r["segments"]) + 1):
                fn = f"run-{r['index']}-boundary-{k}.bin"
                shutil.copyfile(os.path.join(src, "states", fn), os.path.join(dst, fn))
                with open(os.path.join(dst, fn), "rb") as fh:
                    h = state_hash(fh.read())
                want = (r["segments"][k]["start"] if k < len(r["segments"])
                        else r["segments"][-1]["end"])
                assert h == want, (name, fn)
        index[name] = {"system": system, "format": fmt, "runs": runs}
    with open(os.path.join(out_root, "certs.json"), "w") as fh:
        json.dump(index, fh, indent=1, sort_keys=True)
    print(f"wrote {len(index)} certificates under {out_root}")


if __name__ == "__main__":
    main()
