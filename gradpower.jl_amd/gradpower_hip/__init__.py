"""gradpower_hip -- MI355X batched backward-Euler DAE integrator for power-system
transient simulation, a drop-in for GradPower.jl's batched GPU integrator.

Host setup (system model, power flow, device initialisation) is Python; the
integrator itself is the C-ABI library libgradpower_hip.so (HIP kernels for
gfx950), loaded through `gradpower_hip.lib`.
"""
from .system import (Bus, Gen, Load, Branch, Shunt, ContingencyEvent, PowerSystem,
                     PowerSystemDynamics, DynamicProblem, TripLineEvent, create_trip_line_event,
                     DisconnectDeviceEvent, add_trip_event, add_disconnect_event,
                     Genrou, Gensal, IEESGO, TGOV1,
                     SEXS, ESDC1A, IEEEST, ZIPLoad, StaticGenerator, add_device,
                     add_event, set_dynamics, esdc1a_sat_coefficients, PQ, PV, SLACK)
from .network import build_network, runpf, runpf_inplace
from .init import initialize_dynamics
from .desc import build_desc, SystemDesc, gen_speed_indices
from .psse import (from_psse, read_psse_raw, raw_to_grad, read_psse_dyr, create_device_vector,
                   dynamics_from_dyr)

__all__ = [n for n in dir() if not n.startswith("_")]
