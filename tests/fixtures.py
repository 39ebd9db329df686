"""Helpers for the golden-fixture tests: PAM/PNM decoding to canonical RGBA,
and running the reference-CLI test binaries under oracle/_ref/."""
import json
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")


def load_cases():
    with open(os.path.join(GOLDEN, "cases.json")) as f:
        return json.load(f)


def load_expected():
    return np.load(os.path.join(GOLDEN, "expected.npz"))


def read_pnm(data):
    """Decode a binary PNM/PAM byte string -> (maxval, HxWxC uint32)."""
    if data[:3] == b"P7\n":
        hdr_end = data.index(b"ENDHDR\n") + 7
        fields = {}
        for line in data[3:hdr_end].decode().splitlines():
            parts = line.split()
            if len(parts) >= 2:
                fields[parts[0]] = parts[1]
        w, h, c, maxval = (int(fields[k]) for k in ("WIDTH", "HEIGHT", "DEPTH", "MAXVAL"))
        body = data[hdr_end:]
    else:
        kind = data[1:2]
        toks, pos = [], 2
        need = 2 if kind == b"4" else 3
        while len(toks) < need:
            while data[pos:pos + 1].isspace():
                pos += 1
            start = pos
            while not data[pos:pos + 1].isspace():
                pos += 1
            toks.append(int(data[start:pos]))
        pos += 1
        body = data[pos:]
        w, h = toks[0], toks[1]
        if kind == b"4":
            bits = np.unpackbits(np.frombuffer(body, np.uint8)).reshape(h, -1)[:, :w]
            return 1, (1 - bits).astype(np.uint32)[..., None]
        maxval = toks[2]
        c = 3 if kind == b"6" else 1
    dt = ">u2" if maxval > 255 else np.uint8
    arr = np.frombuffer(body, dt, count=w * h * c).reshape(h, w, c).astype(np.uint32)
    return maxval, arr


def canonical(maxval, arr):
    """(H,W,C) samples -> (H,W,4) RGBA at 8 or 16 bits; alpha-0 colour = 0."""
    c = arr.shape[2]
    if c in (1, 2):
        rgb = np.repeat(arr[..., :1], 3, axis=2)
    else:
        rgb = arr[..., :3]
    if c in (2, 4):
        a = arr[..., -1:]
    else:
        a = np.full(arr.shape[:2] + (1,), maxval, np.uint32)
    out = np.concatenate([rgb, a], axis=2)
    depth = 16 if maxval > 255 else 8
    full = (1 << depth) - 1
    if maxval != full:
        out = out * (full // maxval) if full % maxval == 0 else np.round(out * full / maxval).astype(np.uint32)
    out = out.astype(np.uint32)
    out[out[..., 3] == 0, :3] = 0
    return depth, out


def run_cli(binary, case, outdir, timeout=120):
    """Run one case through an oracle/_ref CLI binary; returns (rc, stderr,
    output bytes or None)."""
    src = os.path.join(GOLDEN, "srcimg", case["src"] + ".pam")
    out = os.path.join(outdir, case["name"] + ".pam")
    cmd = [os.path.join(REF_DIR, binary), src, out] + case["args"] + ["-noinfo", "-nowarn"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=timeout)
    data = open(out, "rb").read() if os.path.exists(out) else None
    return r.returncode, r.stderr.decode(errors="replace"), data


def compare_case(binary, case, expected):
    """Returns (ok, message) for one golden case through one binary."""
    with tempfile.TemporaryDirectory() as d:
        rc, err, data = run_cli(binary, case, d)
    if rc != 0 or data is None:
        return False, "exit %d: %s" % (rc, err.strip())
    maxval, arr = read_pnm(data)
    depth, got = canonical(maxval, arr)
    exp = expected[case["name"]].astype(np.uint32)
    if depth == 8 and case["depth"] == 16 and not (exp % 257).any():
        # The PNG writer kept 16 bits only to store a background label
        # exactly; PAM stores no label, so the optimiser reduces to 8 bits.
        exp, depth = exp // 257, 16
    if depth != case["depth"] or got.shape != exp.shape:
        return False, "shape/depth %s/%d vs %s/%d" % (got.shape, depth, exp.shape, case["depth"])
    diff = np.abs(got.astype(np.int64) - exp.astype(np.int64))
    if diff.max() != 0:
        return False, "max diff %d at %d samples" % (diff.max(), int((diff != 0).sum()))
    return True, ""
