"""C ABI and host logic (CPU only: no kernel is launched).

- libmdr_hip.so loads and exports every symbol include/mertens_hip.h declares;
- mdr_plan (pure host logic) gives exactly the oracle's parameters under the
  reference's environment rules;
- option validation and the no-device error path;
- the drop-in header's parameter helpers and environment parsing.
"""
import ctypes
import os
import re
import subprocess

import pytest

import mdr_hip
import oracle

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
HEADER = os.path.join(ROOT, "include", "mertens_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(mdr_[A-Za-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = mdr_hip.load()
    syms = declared_symbols()
    assert len(syms) >= 13
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(mdr_hip.EXPORTS) == syms


def test_struct_sizes_match_header():
    # mdr_opts / mdr_stats layouts as the C compiler sees them
    src = ('#include "mertens_hip.h"\n#include <stdio.h>\n'
           'int main(){printf("%zu %zu\\n", sizeof(mdr_opts), sizeof(mdr_stats));return 0;}\n')
    exe = "/tmp/mdr_sizes"
    subprocess.run(["gcc", "-x", "c", "-I" + os.path.join(ROOT, "include"), "-", "-o", exe], input=src.encode(),
                   check=True)
    a, b = map(int, subprocess.run([exe], capture_output=True, check=True).stdout.split())
    assert a == ctypes.sizeof(mdr_hip.Opts)
    assert b == ctypes.sizeof(mdr_hip.Stats)


def test_default_opts():
    o = mdr_hip.make_opts()
    assert (o.strategy, o.cx, o.u_cap, o.partition, o.params, o.sched, o.chunk) == (0, 0.0, 0, 0, 0, 0, 0)


def visible_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(visible_gpu(), reason="a GPU is visible")
def test_init_without_device_reports_nodevice():
    ctx = ctypes.c_void_p()
    st = mdr_hip.load().mdr_init(ctypes.byref(ctx), 0)
    assert st == 4 and not ctx.value   # MDR_ERR_NODEVICE, nothing allocated


def plan_c(X, **kw):
    lib = mdr_hip.load()
    o = mdr_hip.make_opts(**kw)
    s, u, ud, mu, N, cx = (ctypes.c_int(), ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong(),
                           ctypes.c_longlong(), ctypes.c_double())
    st = lib.mdr_plan(None, X, ctypes.byref(o), *(ctypes.byref(v) for v in (s, u, ud, mu, N, cx)))
    return st, dict(streaming=bool(s.value), u=u.value, u_dense=ud.value, mu_limit=mu.value, N=N.value, cx=cx.value)


ENVS = [{}, {"MERTENS_USE_CHMT": "0"}, {"MERTENS_USE_CHMT": "1"}, {"MERTENS_USE_CHMT": "yes"},
        {"MERTENS_USE_CHMT": "false"}, {"MERTENS_CX": "1.7"}, {"MERTENS_CX": "9"}, {"MERTENS_CX": "abc"},
        {"MERTENS_CHMT_U_CAP": "2000000000"}, {"MERTENS_CHMT_U_CAP": "5"},
        {"MERTENS_FX": "0.5"}, {"MERTENS_FX": "4"}, {"MERTENS_U_CAP": "300000000"},
        {"MERTENS_FX": "0.5", "MERTENS_USE_CHMT": "T"}, {"MERTENS_U_CAP": "900000000", "MERTENS_CX": "0.8"}]
XS = [51000000, 10**8, 987654321, 10**10, 10**11 - 1, 10**11, 123456789012, 10**12, 10**13, 10**14, 10**15,
      10**16, 10**17, 10**18, 9 * 10**18]


@pytest.mark.parametrize("env", ENVS)
def test_plan_matches_oracle(env):
    kw = mdr_hip.opts_from_env(env)
    ok = dict(use_chmt=env.get("MERTENS_USE_CHMT"), cx=env.get("MERTENS_CX"),
              chmt_u_cap=env.get("MERTENS_CHMT_U_CAP"), fx=env.get("MERTENS_FX"), u_cap=env.get("MERTENS_U_CAP"))
    for X in XS:
        st, got = plan_c(X, **kw)
        assert st == 0
        want = oracle.plan(X, **ok)
        assert got == want, (X, env)


def test_invalid_options_rejected():
    for kw in [dict(strategy=7), dict(cx=6.0), dict(u_cap=5), dict(fx=9.0), dict(u_cap_dense=1), dict(partition=3),
               dict(params=2), dict(sched=5), dict(chunk=-1)]:
        st, _ = plan_c(10**12, **kw)
        assert st == 1, kw   # MDR_ERR_INVALID


def test_python_env_parser_follows_c_atof():
    assert mdr_hip._atof("1.5abc") == 1.5 and mdr_hip._atof("x") == 0.0 and mdr_hip._atof(" .5") == 0.5
    assert mdr_hip._atoll("123x") == 123 and mdr_hip._atoll("") == 0 and mdr_hip._atoll("-7") == -7
    assert mdr_hip.opts_from_env({"MERTENS_USE_CHMT": ""}) == {"strategy": mdr_hip.STRATEGY_DENSE}


@pytest.fixture(scope="module")
def plan_probe(tmp_path_factory):
    # built per test process (parallel runners must not overwrite a running copy)
    exe = str(tmp_path_factory.mktemp("plan_probe") / "plan_probe")
    subprocess.run(["g++", "-O1", "-std=c++20", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "plan_probe.cpp"), "-L" + os.path.join(ROOT, "deleglise-rivat_amd"),
                    "-lmdr_hip", "-Wl,-rpath," + os.path.join(ROOT, "deleglise-rivat_amd"), "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("env", ENVS)
def test_dropin_header_parameter_helpers(plan_probe, env):
    full = {k: v for k, v in os.environ.items() if not k.startswith("MERTENS_")}
    full.update(env)
    out = subprocess.run([plan_probe] + [str(x) for x in XS], capture_output=True, text=True, env=full,
                         check=True).stdout.split("\n")
    ok = dict(use_chmt=env.get("MERTENS_USE_CHMT"), cx=env.get("MERTENS_CX"),
              chmt_u_cap=env.get("MERTENS_CHMT_U_CAP"), fx=env.get("MERTENS_FX"), u_cap=env.get("MERTENS_U_CAP"))
    for X, line in zip(XS, out):
        f = line.split()
        assert int(f[0]) == X
        p = oracle.plan(X, **ok)
        assert abs(float(f[1]) - p["cx"]) < 1e-12
        dense = oracle.plan(X, **dict(ok, use_chmt="0"))
        assert int(f[2]) == dense["u"]
        if X >= 10**11:
            stream = oracle.plan(X, **dict(ok, use_chmt="1"))
            assert int(f[3]) == stream["u"]
        assert int(f[6]) == oracle.isqrt(X)
        chmt = env.get("MERTENS_USE_CHMT")
        assert int(f[5]) == (1 if chmt and chmt[:1] in "1yYtT" else 0)
        assert int(f[4]) == (1 if (X >= 10**11 and "MERTENS_FX" not in env and "MERTENS_U_CAP" not in env) else 0)
        # no reference control set: the tuned GPU plan; any set: the reference's rules
        assert int(f[7]) == (mdr_hip.PARAMS_TUNED if not env else mdr_hip.PARAMS_REFERENCE), env


def test_dropin_tuned_switch_overrides_the_default(plan_probe):
    base = {k: v for k, v in os.environ.items() if not k.startswith("MERTENS_")}
    for extra, want in [({"MERTENS_HIP_TUNED": "0"}, mdr_hip.PARAMS_REFERENCE),
                        ({"MERTENS_HIP_TUNED": "1", "MERTENS_CX": "1.7"}, mdr_hip.PARAMS_TUNED),
                        ({"MERTENS_HIP_TUNED": "1"}, mdr_hip.PARAMS_TUNED)]:
        out = subprocess.run([plan_probe, str(10**17)], capture_output=True, text=True, env=dict(base, **extra),
                             check=True).stdout.split()
        assert int(out[7]) == want, extra


def test_cli_usage_and_no_device():
    # the command-line front end (bench_dr / bench_cuda modes): usage text
    # without arguments, a clear error (exit 3) when no GPU is present
    exe = os.path.join(ROOT, "deleglise-rivat_amd", "mdr_bench")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "--benchmark" in r.stdout and "--compare" in r.stdout
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = subprocess.run([exe, "1000"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 3 and "mdr_init failed" in r.stderr
