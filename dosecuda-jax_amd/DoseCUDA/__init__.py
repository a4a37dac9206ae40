"""MI355X-native dose engines (drop-in for DoseCUDA): the proton pencil-beam
and the photon collapsed-cone paths.

``from DoseCUDA import IMPTDoseGrid, IMPTPlan, IMPTBeam`` and
``from DoseCUDA import IMRTDoseGrid, IMRTPlan, IMRTBeam, IMRTControlPoint,
IMRTPhotonEnergy`` work as with the reference package (DoseCUDA/__init__.py:1-4).
"""

from .plan_impt import IMPTDoseGrid, IMPTPlan, IMPTBeam, IMPTBeamModel
from .plan_imrt import IMRTDoseGrid, IMRTPlan, IMRTBeam, IMRTControlPoint, IMRTPhotonEnergy
from .plan import DoseGrid, VolumeObject, Plan, Beam, Prescription

__all__ = ['IMPTDoseGrid', 'IMPTPlan', 'IMPTBeam', 'IMPTBeamModel', 'IMRTDoseGrid', 'IMRTPlan', 'IMRTBeam',
           'IMRTControlPoint', 'IMRTPhotonEnergy', 'DoseGrid', 'VolumeObject', 'Plan', 'Beam', 'Prescription']
