"""Launch sequences of the GPU tests replayed on the CPU simulation build of
the library (tests/cpu_sim): the library's own kernel and engine source runs
on the host with every device allocation fenced by guard pages, so an access
outside its buffer or a barrier that cannot be passed ends the replay
(exit status 86 or 87) before the sequence reaches a GPU.  SIM_GUARD=start
puts the guard before each allocation (underruns), SIM_FILL fills fresh
allocations with garbage as a GPU leaves them."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SIM = os.path.join(HERE, "cpu_sim")


@pytest.fixture(scope="module")
def sim_built():
    # one build at a time (parallel test workers share the output)
    import fcntl
    with open(os.path.join(HERE, ".cpu_sim_build.lock"), "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        subprocess.check_call(["make", "-s", "-C", SIM])


def _run(name, guard="end", fill=None, timeout=600, **kw):
    env = dict(os.environ, SIM_GUARD=guard)
    if fill is not None:
        env["SIM_FILL"] = fill
    args = [f"{k}={v}" for k, v in kw.items()]
    p = subprocess.run([sys.executable, os.path.join(SIM, "scenarios.py"), name] + args, env=env,
                       capture_output=True, text=True, timeout=timeout)
    print(p.stdout[-2000:], p.stderr[-2000:])
    return p


def test_planted_overrun_traps(sim_built):
    """A call for more blocks than its buffers hold must trap in the replay."""
    p = _run("headline_overrun")
    assert p.returncode == 86 and "FAULT" in p.stderr, p.returncode


@pytest.mark.parametrize("guard,fill", [("end", None), ("start", "0xff")])
def test_headline_calls(sim_built, guard, fill):
    p = _run("headline_calls", guard, fill)
    assert p.returncode == 0, p.returncode


def test_streaming_pass(sim_built):
    p = _run("streaming_pass")
    assert p.returncode == 0, p.returncode


@pytest.mark.parametrize("guard,fill", [("end", None), ("start", "0xff")])
def test_c3_group(sim_built, guard, fill):
    p = _run("c3_group", guard, fill)
    assert p.returncode == 0, p.returncode


@pytest.mark.parametrize("guard,fill", [("end", None), ("start", "0xff")])
def test_noalloc_group(sim_built, guard, fill):
    p = _run("noalloc_group", guard, fill)
    assert p.returncode == 0, p.returncode


@pytest.mark.parametrize("guard,fill", [("end", None), ("start", "0xff")])
def test_queue_group(sim_built, guard, fill):
    """Work-queue group launches: a grid of 8 simulated resident workgroups
    takes every launch's items; outputs equal the plain launches'."""
    p = _run("queue_group", guard, fill)
    assert p.returncode == 0, p.returncode


@pytest.mark.parametrize("kp", [4, 8, 16])
def test_pass_phases(sim_built, kp):
    p = _run("pass_phases", "start", "0xff", kp=kp)
    assert p.returncode == 0, p.returncode


@pytest.mark.parametrize("kp", [4, 8, 12, 16])
def test_rt_rig(sim_built, kp):
    p = _run("rt_rig", "start", "0xff", kp=kp)
    assert p.returncode == 0, p.returncode


@pytest.mark.parametrize("kp,B,precision", [(8, 256, 0), (6, 128, 0), (8, 64, 2), (16, 256, 0), (13, 128, 1)])
def test_batched_calls(sim_built, kp, B, precision):
    p = _run("batched_calls", "start", "0xff", kp=kp, B=B, precision=precision)
    assert p.returncode == 0, p.returncode


@pytest.mark.parametrize("log2n", [12, 14, 19])
def test_fft_four_step(sim_built, log2n):
    p = _run("fft_large", "start", "0xff", log2n=log2n)
    assert p.returncode == 0, p.returncode


@pytest.mark.parametrize("guard,fill", [("end", None), ("start", "0xff")])
def test_group_updates(sim_built, guard, fill):
    p = _run("group_updates", guard, fill)
    assert p.returncode == 0, p.returncode
