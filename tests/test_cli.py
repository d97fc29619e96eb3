"""Command-line tools (cu-xmss_amd/ui/xmss_cli.c) with the interface of the
reference's ui/keypair.c, ui/sign.c, ui/open.c.  The CPU tests check the
argument handling (no GPU work happens before it); the GPU tests check that
files written by the tools equal what the C API produces for the same seed,
that sign persists the advanced secret key, and that open accepts exactly
the valid signatures."""
import os
import random
import subprocess

import pytest

from common import product
from ffi import ROOT

BIN = os.path.join(ROOT, "cu-xmss_amd", "bin")
TOOLS = [p + t + c for p in ("xmss_", "xmssmt_") for t in ("keypair", "sign", "open") for c in ("", "_fast")]


def run(tool, *args, env=None, stdin=None):
    e = dict(os.environ)
    if env:
        e.update(env)
    return subprocess.run([os.path.join(BIN, tool)] + list(args), capture_output=True, env=e, timeout=120)


@pytest.mark.parametrize("tool", TOOLS)
def test_tools_built_and_print_usage(tool):
    r = run(tool)
    assert r.returncode != 0
    assert b"usage:" in r.stderr


def test_keypair_rejects_unknown_parameter_set():
    r = run("xmss_keypair", "XMSS-NOPE_10_256")
    assert r.returncode != 0 and b"unknown parameter set" in r.stderr
    r = run("xmssmt_keypair", "XMSS-SHA2_10_256")  # an XMSS name is not an XMSS^MT set
    assert r.returncode != 0


CASES = [("XMSS-SHA2_10_256", 0, 0), ("XMSS-SHA2_10_256", 0, 1), ("XMSSMT-SHA2_20/4_256", 1, 0),
         ("XMSSMT-SHAKE_20/4_256", 1, 1)]


@pytest.mark.gpu
@pytest.mark.parametrize("name,mt,core", CASES)
def test_cli_equals_c_api(tmp_path, name, mt, core):
    api = product(core)
    prefix = ("xmssmt_" if mt else "xmss_")
    sfx = "_fast" if core else ""
    seed = 0xC11 + 7 * mt + core
    print("cli case", name, "core", core, flush=True)
    r = run(prefix + "keypair" + sfx, name, env={"XMSS_MI355X_SEED": hex(seed)})
    assert r.returncode == 0, r.stderr
    oid = api.str_to_oid(name)
    p = api.params(oid, mt)
    api.lib.xmss_mi355x_rng_seed(seed)
    rc, pk, sk = api.keypair(oid, mt, sk_init=bytes(4 + p.sk_bytes))
    assert rc == 0
    assert r.stdout == pk + sk
    kp = tmp_path / "keypair"
    kp.write_bytes(r.stdout)
    rng = random.Random(seed)
    sms = []
    for i in range(3):
        m = bytes(rng.getrandbits(8) for _ in range(rng.randrange(0, 300)))
        mf = tmp_path / ("m%d" % i)
        mf.write_bytes(m)
        r = run(prefix + "sign" + sfx, str(kp), str(mf))
        assert r.returncode == 0, r.stderr
        rc, sm, sk = api.sign(sk, m, mt)
        assert rc == 0
        assert r.stdout == sm, "signature %d" % i
        assert kp.read_bytes() == pk + sk, "persisted sk after signature %d" % i
        sms.append(sm)
    for i, sm in enumerate(sms):
        sf = tmp_path / ("sm%d" % i)
        sf.write_bytes(sm)
        r = run(prefix + "open" + sfx, str(kp), str(sf))
        assert r.returncode == 0 and b"Verification succeeded." in r.stdout
        bad = bytearray(sm)
        bad[p.index_bytes + 5] ^= 1
        sf.write_bytes(bytes(bad))
        r = run(prefix + "open" + sfx, str(kp), str(sf))
        assert r.returncode != 0 and b"Verification failed!" in r.stdout
