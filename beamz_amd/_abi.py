"""ABI constants of ``libbeamz_hip.so``.

Loaded from ``abi_layout.json`` (the single source of truth, mirroring the
reference's ``cuda/abi_layout.json:1-67`` and its generated
``beamz/simulation/_cuda_abi.py:57-118``).  ``include/beamz_hip.h`` restates the
same numbers as ``BZH_*`` macros; a test checks that the three agree.
"""

from __future__ import annotations

import json
import os

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "abi_layout.json")
with open(_PATH, encoding="utf-8") as _fh:
    LAYOUT = json.load(_fh)

ABI_VERSION = int(LAYOUT["abi_version"])
COMPONENT_VERSION = str(LAYOUT["component_version"])
STREAMED_TARGET = LAYOUT["targets"]["streamed"]
PROGRAM_TARGET = LAYOUT["targets"]["program"]

FLAG_GRAPH_CACHE = 1 << int(LAYOUT["flags"]["graph_cache"])
FLAG_BF16_PSI = 1 << int(LAYOUT["flags"]["bf16_psi"])
FLAG_FUSED_STEP = 1 << int(LAYOUT["flags"]["fused_step"])
DEFAULT_FLAGS = sum(1 << int(LAYOUT["flags"][name]) for name in LAYOUT["default_flags"])

_PL = LAYOUT["program_layouts"]
PROGRAM_LAYOUT_YEE_IN_PLACE = int(_PL["yee_in_place"])
PROGRAM_LAYOUT_YEE_TEMPORAL = int(_PL["yee_temporal"])
PROGRAM_LAYOUT_CPML_IN_PLACE = int(_PL["cpml_in_place"])
PROGRAM_LAYOUT_SOURCE_IN_PLACE = int(_PL["source_in_place"])
PROGRAM_LAYOUT_SOURCE_TEMPORAL_CPML = int(_PL["source_temporal_cpml"])
PROGRAM_LAYOUT_MONITOR_IN_PLACE = int(_PL["monitor_in_place"])
PROGRAM_LAYOUT_MONITOR_TEMPORAL_CPML = int(_PL["monitor_temporal_cpml"])

_L = LAYOUT["layout"]
FIELD_COUNT = int(_L["field_count"])
CPML_TERM_COUNT = int(_L["cpml_term_count"])
SOURCE_GROUP_COUNT = int(_L["source_group_count"])
SOURCE_GROUP_BUFFER_COUNT = int(_L["source_group_buffer_count"])
MONITOR_INPUT_COUNT = int(_L["monitor_input_count"])
YEE_GRAPH_INPUT_COUNT = int(_L["yee_graph_input_count"])
CPML_GRAPH_INPUT_COUNT = int(_L["cpml_graph_input_count"])
CPML_PHASE_INPUT_COUNT = int(_L["cpml_phase_input_count"])
CPML_E_INPUT_OFFSET = int(_L["cpml_e_input_offset"])
CPML_H_METRIC_OFFSET = int(_L["cpml_h_metric_offset"])
CPML_E_METRIC_OFFSET = int(_L["cpml_e_metric_offset"])
CPML_H_PSI_INPUT_OFFSET = int(_L["cpml_h_psi_input_offset"])
CPML_E_PSI_INPUT_OFFSET = int(_L["cpml_e_psi_input_offset"])
TEMPORAL_FIELD_WORKSPACE_INPUT = int(_L["temporal_field_workspace_input"])
TEMPORAL_PSI_WORKSPACE_INPUT = int(_L["temporal_psi_workspace_input"])

PHASE_H = 0
PHASE_E = 1
