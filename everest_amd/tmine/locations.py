"""Paths used by the frontend (reference: system/tmine/locations.py:1-60).

Inputs live under $EVEREST_INPUTS (default: <repo>/inputs), with motif files
in motifs/ and graph edge lists in graphs/.  Generated worker sources and
compiled plugins live under everest_amd/tmine/.tmp and everest_amd/plugins.
The label file used for constraint features is $EVEREST_RAND (the reference
hard-codes a path, locations.py:46-47); when unset, labels are generated.
"""
import os


def projRoot():
    return os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def pkgRoot():
    return os.path.join(projRoot(), "everest_amd")


def inputsRoot():
    return os.environ.get("EVEREST_INPUTS", os.path.join(projRoot(), "inputs"))


def inputMotifsRoot():
    return os.path.join(inputsRoot(), "motifs")


def inputGraphRoot():
    return os.path.join(inputsRoot(), "graphs")


def buildDir():
    return os.path.join(pkgRoot(), "lib")


def includeDirs():
    return [os.path.join(projRoot(), "include"), os.path.join(pkgRoot(), "csrc", "hip")]


def systemRoot():
    return os.path.join(pkgRoot(), "tmine")


def templateRoot():
    return os.path.join(systemRoot(), "templates")


def dynLibSourceRoot():
    return os.path.join(systemRoot(), ".tmp")


def dynLibSource():
    return os.path.join(dynLibSourceRoot(), "GPUWorkerDyn.hip")


def dynLibRoot():
    return os.path.join(pkgRoot(), "plugins")


def dynLibPath():
    return os.path.join(dynLibRoot(), "libDYN.so")


def pluginCacheRoot():
    return os.path.join(dynLibRoot(), "cache")


def randFilePath():
    return os.environ.get("EVEREST_RAND", "")


def queryRoot():
    return os.path.join(projRoot(), "inputs", "queries")
