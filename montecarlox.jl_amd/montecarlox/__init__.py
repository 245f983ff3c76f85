"""montecarlox -- MI355X (gfx950) drop-in for the sampling hot path of
MonteCarloX.jl + SpinSystems.

The names mirror the reference's exports (src/MonteCarloX.jl:118-157,
SpinSystems/src/SpinSystems.jl:25-43) with Julia's `!` spelled as a trailing
underscore (`init!` -> `init_`, `update!` -> `update_`).  All sweeps, replica
exchanges and multicanonical / Wang-Landau updates run in libmcx_hip.so; this
package is the host-side object model over its C ABI.
"""
from . import autotune
from ._lib import (Context, MCXArgumentError, MCXAssertionError, MCXBoundsError, MCXError,
                   MCXStateError, default_context)
from .algorithms import (AbstractHeatBath, AbstractImportanceSampling, AbstractMetropolis,
                         Glauber, HeatBath, ImportanceSampling, Metropolis, Multicanonical, Philox,
                         WangLandau, acceptance_rate, ensemble, reset_)
from .batch import LatticeBatch
from .checkpointing import (CheckpointSession, checkpoint_, finalize_, init_checkpoint,
                            restore_checkpoint)
from .ensembles import (BinnedObject, BoltzmannEnsemble, CustomEnsemble, MulticanonicalEnsemble,
                        WangLandauEnsemble, get_centers, get_edges, get_values, logweight,
                        reset_histogram_, set_, update_)
from .muca import (GeneralizedWalkers, ParallelChains, ParallelMulticanonical, ThreadsBackend,
                   distribute_logweight_, merge_histograms_)
from .pt import (ParallelTempering, ReplicaExchange, acceptance_rates, index,
                 optimize_exchange_interval_, set_betas)
from .serving import Job, Server, poisson_workload
from .spinsystems import (BlumeCapel, Ising, IsingLatticeOptim, SpinSystemBatch, energy, init_,
                          magnetization, spin_flip_, sweep_)

__all__ = [n for n in dir() if not n.startswith("_")]
