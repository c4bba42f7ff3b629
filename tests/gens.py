"""Spec -> keys dispatch for the golden set (oracle side and GPU side)."""
import os

from tests.conftest import GOLDEN


def oracle_keys(oracle_mod, r):
    spec = r["spec"]
    g = spec.get("gen")
    if g == "rmat":
        return oracle_mod.gen_rmat(spec["scale"], spec["ef"], spec["a"], spec["b"], spec["c"], spec["seed"])
    if g == "er":
        return oracle_mod.gen_er(spec["n"], spec["samples"], spec["seed"])
    if g == "chung_lu":
        return oracle_mod.gen_chung_lu(spec["n"], spec["gamma"], spec["avg_deg"], spec["seed"])
    if g == "rgg":
        return oracle_mod.gen_rgg(spec["n"], spec["avg_deg"], spec["seed"])
    if g == "sbm":
        return oracle_mod.gen_sbm(spec["blocks"], spec["size"], spec["p_in"], spec["p_out"], spec["seed"])
    st, keys, n, _ = oracle_mod.read_mtx(os.path.join(GOLDEN, spec["file"]))
    assert st == 0
    return keys


def gpu_graph(ctx, r):
    """Builds the golden graph on the GPU (generated in HBM, or read from the fixture file)."""
    import tribit_amd

    spec = r["spec"]
    g = spec.get("gen")
    if g == "rmat":
        return ctx.gen_rmat(spec["scale"], spec["ef"], spec["a"], spec["b"], spec["c"], spec["seed"])
    if g == "er":
        return ctx.gen_er(spec["n"], spec["samples"], spec["seed"])
    if g == "chung_lu":
        return ctx.gen_chung_lu(spec["n"], spec["gamma"], spec["avg_deg"], spec["seed"])
    if g == "rgg":
        return ctx.gen_rgg(spec["n"], spec["avg_deg"], spec["seed"])
    if g == "sbm":
        return ctx.gen_sbm(spec["blocks"], spec["size"], spec["p_in"], spec["p_out"], spec["seed"])
    keys, n, _ = tribit_amd.read_mtx(os.path.join(GOLDEN, spec["file"]))
    return ctx.graph_from_keys(keys, n)
