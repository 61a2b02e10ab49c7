import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import __graft_entry__  # noqa: E402

__graft_entry__.load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libldgpu.so")


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test run without a visible HIP device"
    return torch


_RF_CACHE = {}


@pytest.fixture(scope="session")
def gpu_rf(torch_cuda):
    """RFDecode instances cached per (system, audio)."""
    from lddecode_amd.rfdecode import RFDecode

    def get(system="NTSC", audio=True):
        key = (system, audio)
        if key not in _RF_CACHE:
            _RF_CACHE[key] = RFDecode(system=system, decode_analog_audio=44100 if audio else 0, device=0)
        return _RF_CACHE[key]
    return get


_ORF_CACHE = {}


@pytest.fixture(scope="session")
def oracle_rf():
    from oracle.dsp import OracleRF

    def get(system="NTSC", audio=True):
        key = (system, audio)
        if key not in _ORF_CACHE:
            _ORF_CACHE[key] = OracleRF(system, decode_analog_audio=audio)
        return _ORF_CACHE[key]
    return get
