"""The step ledger that every batch of steps is checked with
(StepLedger.account, src/engine-gpu/time-loop.ts:125-150;
tests/engine-gpu/time-loop.test.ts:55-156 holds its cases), fed
fabricated control readings on the CPU: a set non-physical bit is
NonPhysicalStateError naming the steps between which it happened; a step
count other than the accounted steps plus those submitted is
GpuExecutionError (work the device rejected); the ledger stands still on
an error and moves on a good reading; an unknown batch size skips the
count check."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def readings(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("ledger") / "ledger_check")
    subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", "-Werror", "-o", exe, os.path.join(HERE, "ledger_check.cpp")],
                   check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.splitlines()
    return [(int(line.split(" ", 2)[0]), int(line.split(" ", 2)[1]), line.split(" ", 2)[2] if line.count(" ") > 1
             else "") for line in out]


BS_OK, BS_ERR_NONPHYSICAL, BS_ERR_EXECUTION = 0, 2, 4


def test_good_batches_move_the_ledger(readings):
    assert readings[0][:2] == (BS_OK, 256)
    assert readings[1][:2] == (BS_OK, 512)
    assert readings[2][:2] == (BS_OK, 612)


def test_a_short_batch_is_an_execution_error_and_the_ledger_stands(readings):
    code, accounted, msg = readings[3]
    assert code == BS_ERR_EXECUTION and accounted == 612
    assert "ran 88 of the 256 steps submitted" in msg


def test_the_non_physical_bit_is_reported_first_with_its_steps(readings):
    code, accounted, msg = readings[4]
    assert code == BS_ERR_NONPHYSICAL and accounted == 612
    assert "between steps 612 and 956" in msg


def test_an_unknown_batch_size_skips_the_count(readings):
    assert readings[5][:2] == (BS_OK, 956)
    assert readings[6][:2] == (BS_OK, 1000)
