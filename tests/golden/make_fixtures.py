"""Writes tests/golden/fixture_*.json from the reference's blessed end-to-end fixtures
(tests/oracle/<case>/{recipe,fixture}.json under the reference checkout): the input data of
each case (node coordinates, element connectivity, material, supports, loads) and the blessed
results with their bands. Data only; the tests rebuild the linear system from it with this
repository's own code.

Cases:
  patch-hex8       eight distorted hex8 cells (one mesh.block of one cell each, corners in hex8
                   order; coincident corners are one node), symmetry on x = 0, y = 0, z = 0,
                   tractions 1, 2, 3 on x = 1, y = 1, z = 1; E = 1000, nu = 0.25
  patch-tet10      the same cube as 48 distorted tet10 elements given node by node
  cantilever-tet10 a 10 x 1 x 1 bar, mesh.block tet10 20 x 2 x 2 (Kuhn's split), x = 0 fixed, a
                   total force of 1 in -z split equally over the x = 10 face's nodes; E = 1000,
                   nu = 0 (the mesh is the structured generator's; only parameters are kept)

Usage: python tests/golden/make_fixtures.py [reference checkout, default /root/reference]
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def load(ref, case, name):
    with open(os.path.join(ref, "tests", "oracle", case, name)) as fh:
        return json.load(fh)


def blessed(ref, case):
    return {k: {"value": v["value"], "band": v["band"]} for k, v in load(ref, case, "fixture.json")["values"].items()}


def material(recipe):
    for c in recipe:
        if c["command"] == "model.material.create":
            e = c["params"]["elastic"]
            return float(e["young"]), float(e["poisson"])
    raise ValueError("no material")


def tractions(recipe):
    out = []
    for c in recipe:
        if c["command"] == "model.load.create" and c["params"]["type"] == "traction":
            out.append({"set": c["params"]["set"], "vector": [float(x) for x in c["params"]["vector"]]})
    return out


def patch_hex8(ref):
    recipe = load(ref, "patch-hex8", "recipe.json")
    xyz, conn = [], []

    def node(p):
        for i, q in enumerate(xyz):
            if q == p:
                return i
        xyz.append(p)
        return len(xyz) - 1

    for c in recipe:
        if c["command"] == "mesh.block":
            prm = c["params"]
            assert prm["kind"] == "hex8" and prm["divisions"] == [1, 1, 1]
            conn.append([node([float(v) for v in p]) for p in prm["corners"]])
    young, poisson = material(recipe)
    return {"kind": "hex8", "xyz": xyz, "conn": conn, "young": young, "poisson": poisson,
            "symmetry": ["x", "y", "z"], "tractions": tractions(recipe), "probe_u": [1.0, 1.0, 1.0],
            "probe_stress": [0.46, 0.55, 0.53], "blessed": blessed(ref, "patch-hex8")}


def patch_tet10(ref):
    recipe = load(ref, "patch-tet10", "recipe.json")
    labels, xyz, conn = {}, [], []
    for c in recipe:
        if c["command"] == "mesh.node.add":
            labels[c["params"]["label"]] = len(xyz)
            xyz.append([float(v) for v in c["params"]["xyz"]])
    for c in recipe:
        if c["command"] == "mesh.element.add":
            assert c["params"]["kind"] == "tet10"
            conn.append([labels[v] for v in c["params"]["nodes"]])
    young, poisson = material(recipe)
    return {"kind": "tet10", "xyz": xyz, "conn": conn, "young": young, "poisson": poisson,
            "symmetry": ["x", "y", "z"], "tractions": tractions(recipe), "probe_u": [1.0, 1.0, 1.0],
            "probe_stress": [0.46, 0.55, 0.53], "blessed": blessed(ref, "patch-tet10")}


def cantilever_tet10(ref):
    recipe = load(ref, "cantilever-tet10", "recipe.json")
    blk = next(c["params"] for c in recipe if c["command"] == "mesh.block")
    force = next(c["params"] for c in recipe if c["command"] == "model.load.create")
    young, poisson = material(recipe)
    return {"kind": "tet10", "cells": blk["divisions"], "size": blk["corners"][6], "young": young,
            "poisson": poisson, "clamp": "x0", "force_total": [float(v) for v in force["vector"]],
            "probe_u": [10.0, 0.5, 0.5], "blessed": blessed(ref, "cantilever-tet10")}


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    for name, make in [("patch_hex8", patch_hex8), ("patch_tet10", patch_tet10),
                       ("cantilever_tet10", cantilever_tet10)]:
        with open(os.path.join(HERE, f"fixture_{name}.json"), "w") as fh:
            json.dump(make(ref), fh, indent=1)
        print("wrote", name)


if __name__ == "__main__":
    main()
