"""Command-line tools: flag handling that needs no GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "grus_framework_amd", "bin")


@pytest.fixture(scope="module")
def cli():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "grus_framework_amd", "cli")])
    return BIN


def test_cli_usage_and_app_selection(cli):
    r = subprocess.run([os.path.join(cli, "grus_bfs"), "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--input" in r.stderr and "--pull" in r.stderr
    r = subprocess.run([os.path.join(cli, "grus_bfs"), "--app", "nope"], capture_output=True,
                       text=True)
    assert r.returncode == 2


def test_cli_refuses_multi_gpu(cli):
    r = subprocess.run([os.path.join(cli, "grus_cc"), "--rmat", "8", "--ngpu", "4"],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
