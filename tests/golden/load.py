"""Readers for the committed golden vectors (JSON written by make_golden.py)."""

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load(name):
    with open(os.path.join(HERE, name)) as fh:
        return json.load(fh)


def fhex(values):
    return np.array([float.fromhex(v) for v in values], np.float64)
