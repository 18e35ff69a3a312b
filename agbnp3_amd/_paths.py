"""Locations of the built shared libraries (all built in-tree)."""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_DIR = os.path.join(PKG_DIR, "lib")
CSRC_DIR = os.path.join(PKG_DIR, "csrc")
INCLUDE_DIR = os.path.join(REPO_DIR, "include")

LIBAGBNP3 = os.path.join(LIB_DIR, "libagbnp3.so")
# the same library with the reference's single-precision interface (float_i = float)
LIBAGBNP3_SP = os.path.join(LIB_DIR, "libagbnp3_sp.so")
LIBNBLIST = os.path.join(LIB_DIR, "libnblist.so")
LIBSYNTH = os.path.join(LIB_DIR, "libagbnp3_synth.so")
