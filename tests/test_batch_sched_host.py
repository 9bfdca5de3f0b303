"""Host-only check of the persistent batched kernel's work plans
(gpu_fft_amd/csrc/gfft_batch_sched.h): the native test builds random team
plans and global-level plans, checks their invariants, and simulates the
kernel's loops with threads under 1, 3, 8 and all resident workgroups,
spread over the domains round robin, all in one domain or at random: for team
plans arrival / claim / steal / wait / solo-queue / orphan / reset; for
global-level plans list claims, per-domain done words, the leaders' write-back
and go flags, late steals of absent workgroups' lists, solo transforms during
waits, reset (see tests/native/test_batch_sched.cpp)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_batch_queue_and_simulated_scheduler(tmp_path):
    exe = tmp_path / "test_batch_sched"
    src = os.path.join(ROOT, "tests", "native", "test_batch_sched.cpp")
    inc = os.path.join(ROOT, "gpu_fft_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-pthread", "-fsanitize=address,undefined", "-I", inc, src,
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("ok")
