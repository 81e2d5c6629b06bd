"""Test configuration: import paths and the `gpu` marker."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, 'tf-quant-finance_amd')
for p in (ROOT, PKG):
  if p not in sys.path:
    sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
  config.addinivalue_line('markers', 'gpu: needs a ROCm GPU (MI355X)')


@pytest.fixture(scope='session')
def golden_dir():
  return GOLDEN


def load_golden(name):
  import json
  with open(os.path.join(GOLDEN, name)) as f:
    return json.load(f)
