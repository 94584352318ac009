"""MI355X-native particle step of Gravity (qwertukg/Gravity): P3M gravity,
hard-disc collision substeps with sleeping tiles, position-based liquid,
spatial sort, behind the reference's GravityPipeline API over the C ABI of
libgravity_hip.so; GravityBatch steps many simulations as one."""
from .config import (CollisionConfig, GravityConfig, GravityFieldConfig, LiquidConfig,  # noqa: F401
                     ParticleConfig, StartConfig, TrajectoryConfig)
from .pipeline import (AdaptiveCollisionStatistics, ClusterStatistics, GravityPipeline,  # noqa: F401
                       projectile_offsets)
from .governor import SubstepGovernor  # noqa: F401
from .trajectory import TrajectoryUpdateScheduler  # noqa: F401
from .batch import GravityBatch  # noqa: F401
