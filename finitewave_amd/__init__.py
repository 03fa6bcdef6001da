"""finitewave_amd: the finitewave stepping path on AMD Instinct MI355X.

A drop-in for finitewave's ``model.run()`` path: the same class names and
attributes (tissues, the eight ionic models in 2D and 3D, stencils, stimuli,
trackers, commands, state save/load), with the state held in HBM and each
time step computed by one fused gfx950 kernel (``libfinitewave_amd.so``).

    import finitewave_amd as fw
    tissue = fw.CardiacTissue2D([102, 102])
    model = fw.AlievPanfilov2D()
    ...
    model.run()
"""
from .tissue import CardiacTissue, CardiacTissue2D, CardiacTissue3D
from .stencil import (Stencil, IsotropicStencil2D, IsotropicStencil3D, AsymmetricStencil2D,
                      AsymmetricStencil3D, SymmetricStencil2D)
from .model import CardiacModel
from .models import (AlievPanfilov2D, AlievPanfilov3D, Barkley2D, Barkley3D,
                     MitchellSchaeffer2D, MitchellSchaeffer3D, FentonKarma2D, FentonKarma3D,
                     BuenoOrovio2D, BuenoOrovio3D, LuoRudy912D, LuoRudy913D, TP062D, TP063D,
                     Courtemanche2D, Courtemanche3D)
from .stimulation import (Stim, StimCurrent, StimVoltage, StimSequence, StimCurrentCoord2D,
                          StimCurrentCoord3D, StimVoltageCoord2D, StimVoltageCoord3D,
                          StimCurrentMatrix2D, StimCurrentMatrix3D, StimVoltageMatrix2D,
                          StimVoltageMatrix3D, StimVoltageListMatrix3D, StimCurrentArea2D,
                          StimCurrentArea3D)
from .tracker import (Tracker, TrackerSequence, ActivationTime2DTracker,
                      ActivationTime3DTracker, ActionPotential2DTracker,
                      ActionPotential3DTracker, MultiVariable2DTracker, MultiVariable3DTracker,
                      Variable2DTracker, Variable3DTracker, LocalActivationTime2DTracker,
                      LocalActivationTime3DTracker, Period2DTracker, Period3DTracker,
                      SpiralWaveCore2DTracker, SpiralWaveCore3DTracker, ECG2DTracker,
                      ECG3DTracker)
from .command import Command, CommandSequence
from .fibrosis import (FibrosisPattern, Diffuse2DPattern, Structural2DPattern,
                       Diffuse3DPattern, Structural3DPattern)
from .state import StateSaver, StateSaverCollection, StateLoader
from ._native import NativeUnavailableError

__version__ = "0.1.0"
