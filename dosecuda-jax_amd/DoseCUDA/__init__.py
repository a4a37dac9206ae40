"""MI355X-native proton pencil-beam dose (drop-in for the IMPT part of DoseCUDA).

``from DoseCUDA import IMPTDoseGrid, IMPTPlan, IMPTBeam`` works as with the
reference package (DoseCUDA/__init__.py:1-4); the photon (IMRT) classes of the
reference are not part of this package.
"""

from .plan_impt import IMPTDoseGrid, IMPTPlan, IMPTBeam, IMPTBeamModel
from .plan import DoseGrid, VolumeObject, Plan, Beam, Prescription

__all__ = ['IMPTDoseGrid', 'IMPTPlan', 'IMPTBeam', 'IMPTBeamModel', 'DoseGrid', 'VolumeObject', 'Plan', 'Beam',
           'Prescription']
