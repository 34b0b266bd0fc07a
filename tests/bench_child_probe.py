"""Legs for the benchmark's child-process plumbing test (CPU only)."""
import os


def ok():
    return {"ok": 1}


def abort():
    os.abort()
