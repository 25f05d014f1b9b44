"""The drop-in JSON record of both CLIs: key names and order of the reference's
record (nucleus/cuNucleus34.cu:197-226, 229-797; nucleus/cuNucleus34_ondemand.cu:
338-360, 362-745), its two-level layout (nucleus/utility/json.h:40-92: `{` or
`,` then newline, one tab per level, `"key": value`), and the early exits after
`triangle_count` (no triangles) and after `four_clique_count` (no
four-cliques). The one deliberate key difference is `hip` where the reference
writes `cuda`; `gpu_*` keys are the device's properties."""
import json
import os
import subprocess

import pytest

from conftest import BUILD, GOLDEN

pytestmark = pytest.mark.gpu

ENV = {
    "nucleus34": ["app", "version", "g++", "hip", "parts", "slurm_requested_gpus", "slurm_available_gpus",
                  "slurm_job_gpus", "peer_gpu_count", "thrust", "policy", "procs", "cpu_core_count", "cpu_affinity",
                  "cpu_model", "cpu_physical_cores", "gpu_*", "host"],
    "nucleus34_ondemand": ["app", "version", "g++", "hip", "thrust", "policy", "procs", "cpu_core_count", "cpu_model",
                           "gpu_*", "host"],
}
DATASET = {
    "nucleus34": ["file", "vertex_count", "edge_count", "data_loading_time_sec", "data_memory_bytes",
                  "triangle_time_sec", "triangle_count", "triangles", "|T0", "triangle_graph_time_sec",
                  "four_clique_time_sec", "four_clique_count", "four-cliques", "|Q0", "four_clique_graph_time_sec",
                  "peeling_mode", "peeling_preprocess_time_sec", "peeling_time_sec", "peeling_iterations", "K_max",
                  "total_time_sec", "K_count", "K", "K_avg"],
    "nucleus34_ondemand": ["file", "vertex_count", "edge_count", "data_loading_time_sec", "triangle_time_sec",
                           "triangle_count", "|T0", "incidence_time_sec", "incidence_entries", "four_clique_time_sec",
                           "four_clique_count", "|Q0", "unpeeled_count", "peeling_mode", "peeling_time_sec",
                           "peeling_iterations", "K_max", "total_time_sec", "K_count", "K", "K_avg"],
}


def expected_dataset(exe, case):
    keys = DATASET[exe]
    stop = {"T0": "|T0", "Q0": "|Q0"}.get(case)
    if stop:
        keys = keys[:keys.index(stop)]
    return [k for k in keys if not k.startswith("|")]


def collapse_gpu(keys):
    out = []
    for k in keys:
        k = "gpu_*" if k.startswith("gpu_") else k
        if not (out and out[-1] == k == "gpu_*"):
            out.append(k)
    return out


@pytest.mark.parametrize("exe", ["nucleus34", "nucleus34_ondemand"])
@pytest.mark.parametrize("name,case", [("rmat8.edges", "full"), ("star.edges", "T0"), ("wheel.edges", "Q0")])
def test_json_keys_order_and_layout(tmp_path, exe, name, case):
    print(f"[json] {exe} {name}", flush=True)
    args = [os.path.join(BUILD, exe), os.path.join(GOLDEN, name)]
    r = subprocess.run(args, capture_output=True, text=True, check=True, timeout=120)
    text = r.stdout.strip()
    rec = json.loads(text)
    assert list(rec) == ["environment", "dataset"]
    assert collapse_gpu(list(rec["environment"])) == ENV[exe]
    assert list(rec["dataset"]) == expected_dataset(exe, case), list(rec["dataset"])
    ds = rec["dataset"]
    if case == "T0":
        assert ds["triangle_count"] == 0
    elif case == "Q0":
        assert ds["triangle_count"] > 0 and ds["four_clique_count"] == 0
    else:
        assert ds["K_count"] == ds["triangle_count"] and len(ds["K"]) == min(5, ds["K_count"])
        times = [ds[k] for k in ds if k.endswith("_time_sec") and k != "data_loading_time_sec"]
        assert all(t >= 0 for t in times) and ds["total_time_sec"] > 0
        assert ds["triangle_time_sec"] > 0 and ds["peeling_time_sec"] > 0
        if exe == "nucleus34":
            assert ds["triangle_graph_time_sec"] > 0 and ds["peeling_preprocess_time_sec"] > 0
        else:
            assert ds["incidence_time_sec"] > 0
    # layout: "{", then "\n\t\"environment\":" with the object's "{" opening its
    # first field "\n\t\t\"key\": value", objects closed by "\n\t}" and "\n}"
    lines = text.split("\n")
    assert lines[0] == "{" and lines[1] == '\t"environment":{' and lines[2].startswith('\t\t"app": ')
    assert lines[-2] == "\t}" and lines[-1] == "}"
    assert any(line == '\t},' for line in lines) and any(line == '\t"dataset":{' for line in lines)
    mem = [line.split()[1] for line in r.stderr.splitlines() if line.startswith("[mem] ")]
    if case == "full":
        want = ["graph", "tri-plan", "tri-done", "fc-done", "inc-done"] if exe == "nucleus34" else \
            ["graph", "tri-done", "fc-done"]
        assert mem == want, r.stderr
