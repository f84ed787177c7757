"""cmtj-compatible object model: CVector, drivers, Layer, Junction.

Names, defaults, error types and messages follow the reference's Python
module ``_cmtj`` (python/cmtj.cpp:35-355) so that an existing script runs
unchanged.  Objects are plain Python; ``Junction.runSimulation`` packs them
into the POD descriptors of include/cmtj_amd.h and integrates on the GPU
through ``cmtj_run`` (libcmtj_amd.so).  There is no CPU integration path.
"""
from __future__ import annotations

import copy
import math
import os
import sys
import time
from enum import IntEnum
from typing import Callable, Iterable, Sequence

import numpy as np

from . import _lib

# ---------------------------------------------------------------------------
# enums (python/cmtj.cpp:129-176)


class Axis(IntEnum):
    xaxis = 0
    yaxis = 1
    zaxis = 2
    all = 3
    none = 4


class Reference(IntEnum):
    none = 0
    fixed = 1
    top = 2
    bottom = 3


class SolverMode(IntEnum):
    EulerHeun = 0
    RK4 = 1
    DormandPrince = 2
    Heun = 3


class UpdateType(IntEnum):
    constant = 0
    pulse = 1
    sine = 2
    step = 3
    posine = 4
    halfsine = 5
    trapezoid = 6
    gaussimpulse = 7
    gaussstep = 8
    custom = 100


# ---------------------------------------------------------------------------
# CVector (core/cvector.hpp:13-220, python/cmtj.cpp:101-127)


def _fmt(v: float) -> str:
    # std::ostream << double with default flags (6 significant digits)
    return "%g" % v


class CVector:
    __slots__ = ("x", "y", "z")

    def __init__(self, x: float, y: float, z: float) -> None:
        self.x = float(x)
        self.y = float(y)
        self.z = float(z)

    def length(self) -> float:
        return math.sqrt(self.x * self.x + self.y * self.y + self.z * self.z)

    def normalize(self) -> None:
        mag = self.length()
        if mag != 0:
            self.x = self.x / mag
            self.y = self.y / mag
            self.z = self.z / mag

    def tolist(self) -> list:
        return [self.x, self.y, self.z]

    @staticmethod
    def fromSpherical(theta: float, phi: float, r: float = 1.0) -> "CVector":
        return CVector(r * math.sin(theta) * math.cos(phi), r * math.sin(theta) * math.sin(phi), r * math.cos(theta))

    def __add__(self, o: "CVector") -> "CVector":
        o = _vec(o)
        return CVector(self.x + o.x, self.y + o.y, self.z + o.z)

    def __iadd__(self, o: "CVector") -> "CVector":
        o = _vec(o)
        self.x += o.x
        self.y += o.y
        self.z += o.z
        return self

    def __sub__(self, o: "CVector") -> "CVector":
        o = _vec(o)
        return CVector(self.x - o.x, self.y - o.y, self.z - o.z)

    def __isub__(self, o: "CVector") -> "CVector":
        o = _vec(o)
        self.x -= o.x
        self.y -= o.y
        self.z -= o.z
        return self

    def __mul__(self, v: float) -> "CVector":
        v = float(v)
        return CVector(self.x * v, self.y * v, self.z * v)

    def __rmul__(self, v: float) -> "CVector":
        v = float(v)
        return CVector(v * self.x, v * self.y, v * self.z)

    def __imul__(self, v: float) -> "CVector":
        v = float(v)
        self.x *= v
        self.y *= v
        self.z *= v
        return self

    def __eq__(self, o) -> bool:
        if not isinstance(o, CVector):
            return NotImplemented
        return self.x == o.x and self.y == o.y and self.z == o.z

    def __ne__(self, o) -> bool:
        if not isinstance(o, CVector):
            return NotImplemented
        return not (self.x == o.x and self.y == o.y and self.z == o.z)

    __hash__ = None

    def __getitem__(self, i: int) -> float:
        # operator[] (core/cvector.hpp:161-177): 0 -> x, 1 -> y, else z
        i = int(i)
        if i == 0:
            return self.x
        if i == 1:
            return self.y
        return self.z

    def __iter__(self):
        yield self.x
        yield self.y
        yield self.z

    def __len__(self) -> int:
        return 3

    def __str__(self) -> str:
        return f"[x:{_fmt(self.x)}, y:{_fmt(self.y)}, z:{_fmt(self.z)}]"

    __repr__ = __str__

    def __copy__(self) -> "CVector":
        return CVector(self.x, self.y, self.z)

    def __deepcopy__(self, memo) -> "CVector":
        return CVector(self.x, self.y, self.z)


def _vec(v) -> CVector:
    """Implicit list -> CVector conversion of the binding (python/cmtj.cpp:125-126)."""
    if isinstance(v, CVector):
        return v
    vals = list(v)
    if len(vals) != 3:
        raise RuntimeError("Failed to create vector -- passed list was not of len 3!")
    return CVector(*vals)


def c_dot(a: CVector, b: CVector) -> float:
    """c_dot (core/junction.hpp:82-84)."""
    a, b = _vec(a), _vec(b)
    return a.x * b.x + a.y * b.y + a.z * b.z


# ---------------------------------------------------------------------------
# drivers (core/drivers.hpp:32-461, python/cmtj.cpp:41-98, 178-224)

_DRIVER_FIELDS = (
    "constantValue",
    "amplitude",
    "frequency",
    "phase",
    "period",
    "cycle",
    "timeStart",
    "timeStop",
    "edgeTime",
    "steadyTime",
)


class ScalarDriver:
    """A scalar waveform: value(t) = constantValue + f(t).

    ``ScalarDriver()`` is the constant-zero driver (the binding's only
    constructor); use the static factories for the other waveforms.
    """

    __slots__ = ("update",) + _DRIVER_FIELDS + ("_callback",)

    def __init__(self) -> None:
        self._set(UpdateType.constant)

    def _set(
        self,
        update: UpdateType,
        constantValue: float = 0.0,
        amplitude: float = 0.0,
        frequency: float = -1.0,
        phase: float = 0.0,
        period: float = -1.0,
        cycle: float = -1.0,
        timeStart: float = -1.0,
        timeStop: float = -1.0,
        edgeTime: float = -1.0,
        steadyTime: float = -1.0,
        callback: Callable | None = None,
    ) -> "ScalarDriver":
        # ScalarDriver constructor checks, core/drivers.hpp:111-128
        if update == UpdateType.pulse and (period == -1 or cycle == -1):
            raise RuntimeError(
                "Selected pulse train driver type but either period or cycle were not set"
            )
        if update == UpdateType.sine and frequency == -1:
            raise RuntimeError("Selected sine driver type but frequency was not set")
        self.update = UpdateType(update)
        self.constantValue = float(constantValue)
        self.amplitude = float(amplitude)
        self.frequency = float(frequency)
        self.phase = float(phase)
        self.period = float(period)
        self.cycle = float(cycle)
        self.timeStart = float(timeStart)
        self.timeStop = float(timeStop)
        self.edgeTime = float(edgeTime)
        self.steadyTime = float(steadyTime)
        self._callback = callback
        return self

    @classmethod
    def _make(cls, update, *args, **kw) -> "ScalarDriver":
        d = cls.__new__(cls)
        return d._set(update, *args, **kw)

    # factories, core/drivers.hpp:135-260
    @staticmethod
    def getConstantDriver(constantValue: float) -> "ScalarDriver":
        return ScalarDriver._make(UpdateType.constant, constantValue)

    @staticmethod
    def getPulseDriver(constantValue: float, amplitude: float, period: float, cycle: float) -> "ScalarDriver":
        return ScalarDriver._make(UpdateType.pulse, constantValue, amplitude, -1, -1, period, cycle)

    @staticmethod
    def getSineDriver(constantValue: float, amplitude: float, frequency: float, phase: float) -> "ScalarDriver":
        return ScalarDriver._make(UpdateType.sine, constantValue, amplitude, frequency, phase)

    @staticmethod
    def getPosSineDriver(constantValue: float, amplitude: float, frequency: float, phase: float) -> "ScalarDriver":
        return ScalarDriver._make(UpdateType.posine, constantValue, amplitude, frequency, phase)

    @staticmethod
    def getHalfSineDriver(constantValue: float, amplitude: float, frequency: float, phase: float) -> "ScalarDriver":
        return ScalarDriver._make(UpdateType.halfsine, constantValue, amplitude, frequency, phase)

    @staticmethod
    def getStepDriver(constantValue: float, amplitude: float, timeStart: float, timeStop: float) -> "ScalarDriver":
        if timeStop <= timeStart:
            raise RuntimeError("Start time cannot be later than stop time!")
        return ScalarDriver._make(UpdateType.step, constantValue, amplitude, -1, -1, -1, -1, timeStart, timeStop)

    @staticmethod
    def getTrapezoidDriver(
        constantValue: float, amplitude: float, timeStart: float, edgeTime: float, steadyTime: float
    ) -> "ScalarDriver":
        return ScalarDriver._make(
            UpdateType.trapezoid, constantValue, amplitude, -1, -1, -1, -1, timeStart, -1, edgeTime, steadyTime
        )

    @staticmethod
    def getGaussianImpulseDriver(constantValue: float, amplitude: float, t0: float, sigma: float) -> "ScalarDriver":
        return ScalarDriver._make(UpdateType.gaussimpulse, constantValue, amplitude, -1, -1, -1, -1, t0, -1, sigma)

    @staticmethod
    def getGaussianStepDriver(constantValue: float, amplitude: float, t0: float, sigma: float) -> "ScalarDriver":
        # the reference builds a gaussimpulse here (core/drivers.hpp:241-245)
        return ScalarDriver._make(UpdateType.gaussimpulse, constantValue, amplitude, -1, -1, -1, -1, t0, -1, sigma)

    @staticmethod
    def getCustomDriver(callback: Callable) -> "ScalarDriver":
        # core/drivers.hpp:247-260
        if not callback:
            raise RuntimeError("Callback function is not set")
        code = getattr(callback, "__code__", None)
        if code is None or code.co_argcount != 1:
            raise RuntimeError("Callback function must have one argument")
        return ScalarDriver._make(UpdateType.custom, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, callback=callback)

    # evaluation
    def getCurrentScalarValue(self, time: float) -> float:
        if self.update == UpdateType.custom:
            return _call_custom(self._callback, time)
        return _lib.driver_eval(self._record(), time)

    def _record(self) -> np.ndarray:
        r = np.zeros((), dtype=_lib.DRIVER_DTYPE)
        _fill_driver(r, self)
        return r

    # operators, core/drivers.hpp:313-353 (both scale/shift constant AND amplitude)
    def _derive(self, c: float, a: float) -> "ScalarDriver":
        d = ScalarDriver._make(
            self.update, c, a, self.frequency, self.phase, self.period, self.cycle,
            self.timeStart, self.timeStop, self.edgeTime, self.steadyTime,
        )
        return d

    def __mul__(self, v: float) -> "ScalarDriver":
        return self._derive(self.constantValue * v, self.amplitude * v)

    def __imul__(self, v: float) -> "ScalarDriver":
        self.constantValue *= v
        self.amplitude *= v
        return self

    def __add__(self, v: float) -> "ScalarDriver":
        return self._derive(self.constantValue + v, self.amplitude + v)

    def __iadd__(self, v: float) -> "ScalarDriver":
        self.constantValue += v
        self.amplitude += v
        return self

    def __copy__(self) -> "ScalarDriver":
        d = ScalarDriver.__new__(type(self))
        for f in ScalarDriver.__slots__:
            setattr(d, f, getattr(self, f))
        return d

    def __repr__(self) -> str:
        return f"ScalarDriver({self.update.name}, c={self.constantValue}, A={self.amplitude})"


class NullDriver(ScalarDriver):
    """Constant zero (core/drivers.hpp:356-360)."""

    __slots__ = ()

    def __init__(self) -> None:
        super().__init__()

    def getCurrentScalarValue(self, time: float) -> float:
        return 0.0


def _as_scalar_driver(d) -> ScalarDriver:
    if isinstance(d, ScalarDriver):
        # stored by value: a NullDriver slices to the constant-zero driver
        c = copy.copy(d)
        if type(c) is not ScalarDriver:
            c.__class__ = ScalarDriver
        return c
    raise TypeError(f"expected a ScalarDriver, got {type(d).__name__}")


def _call_custom(cb, time: float) -> float:
    try:
        return float(cb(time))
    except Exception as e:  # drivers.hpp:294-302
        sys.stderr.write(f"Error in Python callback: {e}\n")
        raise RuntimeError("Error in Python callback") from e


class AxialDriver:
    """Three scalar drivers (x, y, z) (core/drivers.hpp:362-461)."""

    __slots__ = ("drivers",)

    def __init__(self, *args, **kwargs) -> None:
        if kwargs:
            args = args + tuple(kwargs.values())
        if len(args) == 3 and all(isinstance(a, ScalarDriver) for a in args):
            self.drivers = [_as_scalar_driver(a) for a in args]
        elif len(args) == 1 and isinstance(args[0], CVector):
            v = args[0]
            self.drivers = [ScalarDriver.getConstantDriver(c) for c in (v.x, v.y, v.z)]
        elif len(args) == 1 and isinstance(args[0], (list, tuple)) and all(isinstance(a, ScalarDriver) for a in args[0]):
            if len(args[0]) != 3:
                raise RuntimeError("The axial driver can only have 3 axes!")
            self.drivers = [_as_scalar_driver(a) for a in args[0]]
        elif len(args) == 3:
            self.drivers = [ScalarDriver.getConstantDriver(float(a)) for a in args]
        elif len(args) == 1:
            v = _vec(args[0])
            self.drivers = [ScalarDriver.getConstantDriver(c) for c in (v.x, v.y, v.z)]
        elif len(args) == 0:
            self.drivers = [ScalarDriver(), ScalarDriver(), ScalarDriver()]
        else:
            raise TypeError("AxialDriver(): incompatible constructor arguments")

    @staticmethod
    def getVectorAxialDriver(x: float, y: float, z: float) -> "AxialDriver":
        return AxialDriver(CVector(x, y, z))

    def getCurrentAxialDrivers(self, time: float) -> CVector:
        return CVector(*(d.getCurrentScalarValue(time) for d in self.drivers))

    def applyMask(self, mask) -> None:
        # core/drivers.hpp:371-387
        if isinstance(mask, CVector):
            vals = [int(mask[0]), int(mask[1]), int(mask[2])]
        else:
            vals = [int(m) for m in mask]
        assert len(vals) == 3
        for i, v in enumerate(vals):
            if v == 0:
                self.drivers[i] = ScalarDriver()
            elif v != 1:
                raise RuntimeError("Invalid mask value, mask must be binary!")

    def __copy__(self) -> "AxialDriver":
        a = AxialDriver.__new__(AxialDriver)
        a.drivers = [copy.copy(d) for d in self.drivers]
        return a


def _as_axial(d) -> AxialDriver:
    if isinstance(d, AxialDriver):
        return copy.copy(d)
    raise TypeError(f"expected an AxialDriver, got {type(d).__name__}")


# module-level driver aliases, python/cmtj.cpp:41-98
def constantDriver(value: float) -> ScalarDriver:
    return ScalarDriver.getConstantDriver(value)


def pulseDriver(constantValue: float, amplitude: float, period: float, cycle: float) -> ScalarDriver:
    return ScalarDriver.getPulseDriver(constantValue, amplitude, period, cycle)


def sineDriver(constantValue: float, amplitude: float, frequency: float, phase: float) -> ScalarDriver:
    return ScalarDriver.getSineDriver(constantValue, amplitude, frequency, phase)


def posSineDriver(constantValue: float, amplitude: float, frequency: float, phase: float) -> ScalarDriver:
    return ScalarDriver.getPosSineDriver(constantValue, amplitude, frequency, phase)


def stepDriver(constantValue: float, amplitude: float, timeStart: float, timeStop: float) -> ScalarDriver:
    return ScalarDriver.getStepDriver(constantValue, amplitude, timeStart, timeStop)


def trapezoidDriver(
    constantValue: float, amplitude: float, timeStart: float, edgeTime: float, steadyTime: float
) -> ScalarDriver:
    return ScalarDriver.getTrapezoidDriver(constantValue, amplitude, timeStart, edgeTime, steadyTime)


def gaussianImpulseDriver(constantValue: float, amplitude: float, t0: float, sigma: float) -> ScalarDriver:
    return ScalarDriver.getGaussianImpulseDriver(constantValue, amplitude, t0, sigma)


def gaussianStepDriver(constantValue: float, amplitude: float, t0: float, sigma: float) -> ScalarDriver:
    return ScalarDriver.getGaussianStepDriver(constantValue, amplitude, t0, sigma)


def _driver_row(d: ScalarDriver, table_index: int = -1) -> tuple:
    """One cmtj_driver record as a tuple in field order."""
    if d.update == UpdateType.custom:
        return (_lib.TABLE, table_index, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    return (int(d.update), 0, d.constantValue, d.amplitude, d.frequency, d.phase, d.period, d.cycle, d.timeStart,
            d.timeStop, d.edgeTime, d.steadyTime)


def _fill_driver(rec: np.ndarray, d: ScalarDriver, table_index: int = -1) -> None:
    rec[()] = _driver_row(d, table_index)


# ---------------------------------------------------------------------------
# AdaptiveIntegrationParams (core/junction.hpp:125-139, python/cmtj.cpp:151-163)

class AdaptiveIntegrationParams:
    """Step control of the Dormand-Prince solver; defaults of the reference."""

    def __init__(self) -> None:
        self.abs_tol = 1e-6
        self.rel_tol = 1e-3
        self.max_factor = 5.0
        self.min_factor = 0.1
        self.safety_factor = 0.9
        self.use_pid_control = False
        self.ki = 0.0
        self.kp = 0.2
        self.kd = 0.0
        self.prev_error_ratio = 1.0
        self.integral_error = 0.0

    def __copy__(self) -> "AdaptiveIntegrationParams":
        p = AdaptiveIntegrationParams()
        p.__dict__.update(self.__dict__)
        return p


# ---------------------------------------------------------------------------
# Layer (core/junction.hpp:145-1137, python/cmtj.cpp:226-289)

_SCALAR_SLOTS = {
    "anisotropy": _lib.SLOT_K,
    "secondOrderAnisotropy": _lib.SLOT_K2,
    "iecTop": _lib.SLOT_J_TOP,
    "iecBottom": _lib.SLOT_J_BOTTOM,
    "quadIecTop": _lib.SLOT_J2_TOP,
    "quadIecBottom": _lib.SLOT_J2_BOTTOM,
    "current": _lib.SLOT_CURRENT,
    "fieldLike": _lib.SLOT_FL,
    "dampingLike": _lib.SLOT_DL,
    "fieldLike2": _lib.SLOT_FL2,
    "dampingLike2": _lib.SLOT_DL2,
    "temperature": _lib.SLOT_TEMPERATURE,
}
_AXIAL_SLOTS = {
    "external": _lib.SLOT_HEXT,
    "oersted": _lib.SLOT_HOE,
    "hdmi": _lib.SLOT_HDMI,
    "reserved": _lib.SLOT_HRES,
    "idmiTop": _lib.SLOT_IDMI_TOP,
    "idmiBottom": _lib.SLOT_IDMI_BOTTOM,
}


def _tensor(rows) -> list:
    rows = [_vec(r) for r in rows]
    if len(rows) != 3:
        raise RuntimeError("tensor must have 3 rows")
    return rows


class Layer:
    """One magnetic layer (reference Layer<T>)."""

    def __init__(
        self,
        id: str,
        mag: CVector,
        anis: CVector,
        Ms: float,
        thickness: float,
        cellSurface: float,
        demagTensor: Sequence[CVector],
        damping: float = 0.011,
    ) -> None:
        self._init(id, mag, anis, Ms, thickness, cellSurface, demagTensor, damping, 0.0, 0.0, 0.0, 0.0, 0.0)

    def _init(self, id, mag, anis, Ms, thickness, cellSurface, demagTensor, damping,
              fieldLikeTorque, dampingLikeTorque, slonczewski, beta, spinPolarisation) -> None:
        # private constructor, core/junction.hpp:178-200
        mag = copy.copy(_vec(mag))
        anis = copy.copy(_vec(anis))
        if mag.length() == 0:
            raise RuntimeError("Initial magnetisation was set to a zero vector!")
        if anis.length() == 0:
            raise RuntimeError("Anisotropy was set to a zero vector!")
        self.id = str(id)
        # the stored magnetisation is NOT normalised (the reference normalises
        # its by-value parameter only, core/junction.hpp:195)
        self.mag = mag
        self.anis = anis
        self.Ms = float(Ms)
        self.thickness = float(thickness)
        self.cellSurface = float(cellSurface)
        self.cellVolume = self.cellSurface * self.thickness
        self.demagTensor = [copy.copy(v) for v in _tensor(demagTensor)]
        self.damping = float(damping)
        self.fieldLikeTorque = float(fieldLikeTorque)
        self.dampingLikeTorque = float(dampingLikeTorque)
        self.SlonczewskiSpacerLayerParameter = float(slonczewski)
        self.beta = float(beta)
        self.spinPolarisation = float(spinPolarisation)
        self.kappa = 1.0
        self.includeSTT = False
        self.includeSOT = False
        self.dynamicSOT = True
        self.alternativeSTT = False
        self.temperatureSet = False
        self.nonStochasticTempSet = False
        self.referenceType = Reference.none
        self.referenceLayer = CVector(0, 0, 0)
        self.secondaryReferenceLayer = CVector(0, 0, 0)
        self.dipoleTop = [CVector(0, 0, 0) for _ in range(3)]
        self.dipoleBottom = [CVector(0, 0, 0) for _ in range(3)]
        self.alphaNoise = 1.0
        self.stdNoise = 0.0
        self.scaleNoise = 0.0
        self.noiseAxis = Axis.all
        self.oneFNoise = None
        self.adaptiveParams = AdaptiveIntegrationParams()
        self._scalar = {k: ScalarDriver() for k in _SCALAR_SLOTS}
        self._axial = {k: AxialDriver() for k in _AXIAL_SLOTS}

    # factories, core/junction.hpp:334-353 / python/cmtj.cpp:238-246
    @staticmethod
    def createSOTLayer(
        id: str, mag: CVector, anis: CVector, Ms: float, thickness: float, cellSurface: float,
        demagTensor: Sequence[CVector], damping: float = 0.011, fieldLikeTorque: float = 0.0,
        dampingLikeTorque: float = 0.0,
    ) -> "Layer":
        L = Layer.__new__(Layer)
        L._init(id, mag, anis, Ms, thickness, cellSurface, demagTensor, damping,
                fieldLikeTorque, dampingLikeTorque, 0.0, 0.0, 0.0)
        L.includeSTT = False
        L.includeSOT = True
        L.dynamicSOT = False
        return L

    @staticmethod
    def createSTTLayer(
        id: str, mag: CVector, anis: CVector, Ms: float, thickness: float, cellSurface: float,
        demagTensor: Sequence[CVector], damping: float = 0.011, SlonczewskiSpacerLayerParameter: float = 1.0,
        beta: float = 0.0, spinPolarisation: float = 0.0,
    ) -> "Layer":
        L = Layer.__new__(Layer)
        L._init(id, mag, anis, Ms, thickness, cellSurface, demagTensor, damping,
                0.0, 0.0, SlonczewskiSpacerLayerParameter, beta, spinPolarisation)
        L.includeSTT = True
        L.includeSOT = False
        return L

    def getId(self) -> str:
        return self.id

    def setMagnetisation(self, mag: CVector) -> None:
        mag = copy.copy(_vec(mag))
        if mag.length() == 0:
            raise RuntimeError("Initial magnetisation was set to a zero vector!")
        mag.normalize()
        self.mag = mag

    def setAnisotropyDriver(self, driver: ScalarDriver) -> None:
        self._scalar["anisotropy"] = _as_scalar_driver(driver)

    def setSecondOrderAnisotropyDriver(self, driver: ScalarDriver) -> None:
        self._scalar["secondOrderAnisotropy"] = _as_scalar_driver(driver)

    def setExternalFieldDriver(self, driver: AxialDriver) -> None:
        self._axial["external"] = _as_axial(driver)

    def setOerstedFieldDriver(self, driver: AxialDriver) -> None:
        self._axial["oersted"] = _as_axial(driver)

    def setHdmiDriver(self, driver: AxialDriver) -> None:
        self._axial["hdmi"] = _as_axial(driver)

    def setReservedInteractionField(self, driver: AxialDriver) -> None:
        self._axial["reserved"] = _as_axial(driver)

    def setCurrentDriver(self, driver: ScalarDriver) -> None:
        self._scalar["current"] = _as_scalar_driver(driver)

    def setTemperatureDriver(self, driver: ScalarDriver) -> None:
        self._scalar["temperature"] = _as_scalar_driver(driver)
        self.temperatureSet = True

    def setNonStochasticLangevinDriver(self, driver: ScalarDriver) -> None:
        self._scalar["temperature"] = _as_scalar_driver(driver)
        self.temperatureSet = False
        self.nonStochasticTempSet = True

    def _setTorqueParameters(self) -> None:
        # core/junction.hpp:431-439
        self.includeSOT = True
        if self.includeSTT:
            raise RuntimeError("includeSTT was on and now setting SOT interaction!")
        if not self.dynamicSOT:
            raise RuntimeError("used a static SOT definition, now trying to set it dynamically!")

    def setFieldLikeTorqueDriver(self, driver: ScalarDriver) -> None:
        self._setTorqueParameters()
        self._scalar["fieldLike"] = _as_scalar_driver(driver)

    def setDampingLikeTorqueDriver(self, driver: ScalarDriver) -> None:
        self._setTorqueParameters()
        self._scalar["dampingLike"] = _as_scalar_driver(driver)

    def setSecondaryFieldLikeTorqueDriver(self, driver: ScalarDriver) -> None:
        self._setTorqueParameters()
        self._scalar["fieldLike2"] = _as_scalar_driver(driver)

    def setSecondaryDampingLikeTorqueDriver(self, driver: ScalarDriver) -> None:
        self._setTorqueParameters()
        self._scalar["dampingLike2"] = _as_scalar_driver(driver)

    def setPrimaryTorqueDrivers(self, fieldLikeTorque: ScalarDriver, dampingLikeTorque: ScalarDriver) -> None:
        self._setTorqueParameters()
        self._scalar["fieldLike"] = _as_scalar_driver(fieldLikeTorque)
        self._scalar["dampingLike"] = _as_scalar_driver(dampingLikeTorque)

    def setSecondaryTorqueDrivers(self, fieldLikeTorque: ScalarDriver, dampingLikeTorque: ScalarDriver) -> None:
        self._setTorqueParameters()
        self._scalar["fieldLike2"] = _as_scalar_driver(fieldLikeTorque)
        self._scalar["dampingLike2"] = _as_scalar_driver(dampingLikeTorque)

    def setIECDriverTop(self, driver: ScalarDriver) -> None:
        self._scalar["iecTop"] = _as_scalar_driver(driver)

    def setIECDriverBottom(self, driver: ScalarDriver) -> None:
        self._scalar["iecBottom"] = _as_scalar_driver(driver)

    def setQuadIECDriverTop(self, driver: ScalarDriver) -> None:
        self._scalar["quadIecTop"] = _as_scalar_driver(driver)

    def setQuadIECDriverBottom(self, driver: ScalarDriver) -> None:
        self._scalar["quadIecBottom"] = _as_scalar_driver(driver)

    def setIDMIDriverTop(self, driver: AxialDriver) -> None:
        self._axial["idmiTop"] = _as_axial(driver)

    def setIDMIDriverBottom(self, driver: AxialDriver) -> None:
        self._axial["idmiBottom"] = _as_axial(driver)

    def setReferenceLayer(self, ref) -> None:
        # two overloads, core/junction.hpp:517-548
        if isinstance(ref, Reference):
            if ref == Reference.fixed and not self.referenceLayer.length():
                raise RuntimeError(
                    "Cannot set fixed polarisation layer to 0! Set reference to NONE to disable reference."
                )
            self.referenceType = ref
            return
        self.referenceLayer = copy.copy(_vec(ref))
        self.referenceType = Reference.fixed

    def setSecondaryReferenceLayer(self, ref: CVector) -> None:
        self.secondaryReferenceLayer = copy.copy(_vec(ref))
        self.referenceType = Reference.fixed

    def getReferenceLayer(self) -> CVector:
        return self.referenceLayer

    def getSecondaryReferenceLayer(self) -> CVector:
        return self.secondaryReferenceLayer

    def getReferenceType(self) -> Reference:
        return self.referenceType

    def setTopDipoleTensor(self, tensor: Sequence[CVector]) -> None:
        self.dipoleTop = [copy.copy(v) for v in _tensor(tensor)]

    def setBottomDipoleTensor(self, tensor: Sequence[CVector]) -> None:
        self.dipoleBottom = [copy.copy(v) for v in _tensor(tensor)]

    def setKappa(self, kappa: float) -> None:
        self.kappa = float(kappa)

    def setAlternativeSTT(self, setAlternative: bool) -> None:
        self.alternativeSTT = bool(setAlternative)

    def hasTemperature(self) -> bool:
        return self.temperatureSet

    # coloured (1/f, alpha) noise belongs to core/noise.hpp, which this engine
    # does not provide; the setters record the request and a run refuses it
    def setAlphaNoise(self, alpha: float, std: float, scale: float, axis: Axis = Axis.all) -> None:
        if alpha < 0 or alpha > 2:
            raise RuntimeError("alpha must be between 0 and 2")
        self.alphaNoise, self.stdNoise, self.scaleNoise, self.noiseAxis = float(alpha), float(std), float(scale), axis

    def setOneFNoise(self, sources: int, bias: float, scale: float) -> None:
        self.oneFNoise = (int(sources), float(bias), float(scale))

    def createBufferedAlphaNoise(self, bufferSize: int) -> None:
        raise NotImplementedError("coloured (alpha) noise is not provided by the GPU engine")

    def getOneFVector(self) -> CVector:
        if self.scaleNoise != 0:
            raise NotImplementedError("coloured (alpha) noise is not provided by the GPU engine")
        return CVector(0, 0, 0)

    def setAdaptiveParams(self, params) -> None:
        # the layer keeps its own copy (core/junction.hpp:355-357)
        self.adaptiveParams = copy.copy(params)

    def __copy__(self) -> "Layer":
        L = Layer.__new__(Layer)
        for k, v in self.__dict__.items():
            setattr(L, k, copy.deepcopy(v) if k not in ("_scalar", "_axial") else None)
        L._scalar = {k: copy.copy(v) for k, v in self._scalar.items()}
        L._axial = {k: copy.copy(v) for k, v in self._axial.items()}
        return L

    def __repr__(self) -> str:
        return f"Layer(id={self.id!r}, Ms={self.Ms}, thickness={self.thickness})"

    # descriptor ------------------------------------------------------------
    def _fill(self, tables: "_Tables", run_info) -> tuple:
        """The layer's cmtj_layer record as one tuple in the dtype's field
        order (assigned with one numpy call instead of one per field)."""
        rows = [None] * _lib.NSLOT
        for name, slot in _SCALAR_SLOTS.items():
            d = self._scalar[name]
            rows[slot] = _driver_row(d, tables.add(d, run_info) if d.update == UpdateType.custom else -1)
        for name, slot in _AXIAL_SLOTS.items():
            for q, d in enumerate(self._axial[name].drivers):
                rows[slot + q] = _driver_row(d, tables.add(d, run_info) if d.update == UpdateType.custom else -1)
        ap = self.adaptiveParams
        return (
            (self.mag.x, self.mag.y, self.mag.z),
            (self.anis.x, self.anis.y, self.anis.z),
            self.Ms, self.thickness, self.cellSurface, self.damping,
            tuple(c for v in self.demagTensor for c in (v.x, v.y, v.z)),
            tuple(c for v in self.dipoleTop for c in (v.x, v.y, v.z)),
            tuple(c for v in self.dipoleBottom for c in (v.x, v.y, v.z)),
            tuple(self.referenceLayer.tolist()),
            tuple(self.secondaryReferenceLayer.tolist()),
            self.kappa, self.beta, self.SlonczewskiSpacerLayerParameter, self.spinPolarisation,
            self.fieldLikeTorque, self.dampingLikeTorque,
            int(self.referenceType), int(self.includeSTT), int(self.includeSOT), int(self.dynamicSOT),
            int(self.alternativeSTT), int(self.temperatureSet),
            rows,
            ap.abs_tol, ap.rel_tol, ap.max_factor, ap.min_factor, ap.safety_factor,
            int(bool(ap.use_pid_control)), 0, ap.kp, ap.ki, ap.kd, ap.integral_error, ap.prev_error_ratio,
        )


class _Tables:
    """Samples of Python-callback drivers at the solver's stage times.

    The reference calls the callback at every stage (core/drivers.hpp:
    294-302); the same calls are made here before the launch, at the same
    binary64 times, and the kernel reads the samples."""

    def __init__(self) -> None:
        self.chunks: list[np.ndarray] = []
        self.size = 0

    def add(self, d: ScalarDriver, run_info) -> int:
        n_steps, dt, solver = run_info
        i = np.arange(n_steps, dtype=np.float64)
        t = i * dt
        if solver == _lib.RK4:
            times = np.stack([t, t + 0.5 * dt, t + dt], axis=1)
        elif solver == _lib.HEUN:
            times = np.stack([t, t + dt], axis=1)
        else:
            times = t[:, None]
        cb = d._callback
        vals = np.fromiter((_call_custom(cb, float(x)) for x in times.ravel()), dtype=np.float64, count=times.size)
        idx = self.size
        self.chunks.append(vals)
        self.size += vals.size
        return idx

    def array(self) -> np.ndarray | None:
        if not self.chunks:
            return None
        return np.concatenate(self.chunks)


# ---------------------------------------------------------------------------
# Junction (core/junction.hpp:1139-2057, python/cmtj.cpp:291-355)

_ENERGY_KEYS = ("K", "K2", "Jbottom", "Jtop", "I", "Hextx", "Hexty", "Hextz", "Hiecx", "Hiecy", "Hiecz",
                "Hanisx", "Hanisy", "Hanisz", "Hdemagx", "Hdemagy", "Hdemagz", "Hflx", "Hfly", "Hflz",
                "Hdlx", "Hdly", "Hdlz")


class Junction:
    """A stack of layers integrated together (reference Junction<T>)."""

    def __init__(self, *args, **kwargs) -> None:
        names = ("layers", "Rx0", "Ry0", "AMR_X", "AMR_Y", "SMR_X", "SMR_Y", "AHE")
        if len(args) + len(kwargs) == 8 or any(k in kwargs for k in names[1:]):
            vals = dict(zip(names, args))
            vals.update(kwargs)
            self._init_strip(*(vals[k] for k in names))
            return
        layers = kwargs.pop("layers", args[0] if args else None)
        rest = args[1:]
        if layers is None:
            raise TypeError("Junction(): missing layers")
        if not rest and not kwargs:
            self._init_plain(layers)
            return
        Rp = kwargs.pop("Rp", rest[0] if len(rest) > 0 else 100.0)
        Rap = kwargs.pop("Rap", rest[1] if len(rest) > 1 else 200.0)
        self._init_classic(layers, Rp, Rap)

    def _common(self, layers) -> None:
        self.layers = [copy.copy(l) for l in layers]
        self.layerNo = len(self.layers)
        self.MR_mode = _lib.MR_NONE
        self.Rp = 0.0
        self.Rap = 0.0
        self.Rx0 = self.Ry0 = self.AMR_X = self.AMR_Y = self.SMR_X = self.SMR_Y = self.AHE = []
        self.Rtag = "R"
        self._log: dict[str, list] = {}
        self.logLength = 0
        self._seed = int.from_bytes(os.urandom(8), "little")
        self._runs = 0

    def __deepcopy__(self, memo) -> "Junction":
        """A copy that owns its layers, drivers and log, as the reference's
        containers copy junctions by value; layer by layer instead of a
        generic deep copy, which dominates building many stacks or groups."""
        J = self.__class__.__new__(self.__class__)
        memo[id(self)] = J
        for k, v in self.__dict__.items():
            if k == "layers":
                v = [copy.copy(l) for l in v]
            elif k == "_log":  # chunks are never written in place
                v = {key: list(chunks) for key, chunks in v.items()}
            else:
                v = copy.deepcopy(v, memo)
            J.__dict__[k] = v
        return J

    def _init_plain(self, layers) -> None:
        # core/junction.hpp:1164-1180
        self._common(layers)
        if self.layerNo == 0:
            raise ValueError("Passed a zero length Layer vector!")
        ids = set()
        for l in self.layers:
            if l.id in ids:
                raise ValueError("Layers must have unique ids!")
            ids.add(l.id)

    def _init_classic(self, layers, Rp, Rap) -> None:
        # core/junction.hpp:1181-1203
        self._init_plain(layers)
        if self.layerNo == 1 and not self.layers[0].referenceLayer.length():
            raise ValueError("MTJ with a single layer must have a pinning (referenceLayer) set!")
        if self.layerNo > 2:
            raise ValueError(
                "This constructor supports only bilayers! Choose the other one with the strip resistance!"
            )
        self.Rp = float(Rp)
        self.Rap = float(Rap)
        self.MR_mode = _lib.MR_CLASSIC
        if self.layerNo == 2:
            self.Rtag = "R_" + self.layers[0].id + "_" + self.layers[1].id

    def _init_strip(self, layers, Rx0, Ry0, AMR_X, AMR_Y, SMR_X, SMR_Y, AHE) -> None:
        # core/junction.hpp:1219-1243
        self._common(layers)
        if self.layerNo == 0:
            raise ValueError("Passed a zero length Layer vector!")
        arrs = [list(map(float, a)) for a in (Rx0, Ry0, AMR_X, AMR_Y, SMR_X, SMR_Y, AHE)]
        if any(len(a) != self.layerNo for a in arrs):
            raise ValueError("Layers and Rx0, Ry, AMR, AMR and SMR must be of the same size!")
        self.Rx0, self.Ry0, self.AMR_X, self.AMR_Y, self.SMR_X, self.SMR_Y, self.AHE = arrs
        self.MR_mode = _lib.MR_STRIP

    # --- layer access -------------------------------------------------------
    def getLayerIds(self) -> list:
        return [l.id for l in self.layers]

    def getLayer(self, layerId: str) -> Layer:
        for l in self.layers:
            if l.id == layerId:
                return l
        raise RuntimeError("Failed to find a layer with a given id " + layerId + "!")

    def _each(self, layerID: str) -> list:
        found = [l for l in self.layers if l.id == layerID or layerID == "all"]
        if not found:
            raise RuntimeError("Failed to find a layer with a given id: " + layerID + "!")
        return found

    def _scalar_set(self, layerID: str, method: str, driver) -> None:
        for l in self._each(layerID):
            getattr(l, method)(driver)

    def setLayerTemperatureDriver(self, layerId: str, driver: ScalarDriver) -> None:
        self._scalar_set(layerId, "setTemperatureDriver", driver)

    def setLayerNonStochasticLangevinDriver(self, layerId: str, driver: ScalarDriver) -> None:
        self._scalar_set(layerId, "setNonStochasticLangevinDriver", driver)

    def setLayerAnisotropyDriver(self, layerId: str, driver: ScalarDriver) -> None:
        self._scalar_set(layerId, "setAnisotropyDriver", driver)

    def setLayerSecondOrderAnisotropyDriver(self, layerId: str, driver: ScalarDriver) -> None:
        self._scalar_set(layerId, "setSecondOrderAnisotropyDriver", driver)

    def setLayerExternalFieldDriver(self, layerId: str, driver: AxialDriver) -> None:
        self._scalar_set(layerId, "setExternalFieldDriver", driver)

    def setLayerOerstedFieldDriver(self, layerId: str, driver: AxialDriver) -> None:
        self._scalar_set(layerId, "setOerstedFieldDriver", driver)

    def setLayerHdmiDriver(self, layerId: str, driver: AxialDriver) -> None:
        self._scalar_set(layerId, "setHdmiDriver", driver)

    def setLayerCurrentDriver(self, layerId: str, driver: ScalarDriver) -> None:
        self._scalar_set(layerId, "setCurrentDriver", driver)

    def setLayerDampingLikeTorqueDriver(self, layerId: str, driver: ScalarDriver) -> None:
        self._scalar_set(layerId, "setDampingLikeTorqueDriver", driver)

    def setLayerFieldLikeTorqueDriver(self, layerId: str, driver: ScalarDriver) -> None:
        self._scalar_set(layerId, "setFieldLikeTorqueDriver", driver)

    def setLayerSecondaryFieldLikeTorqueDriver(self, layerId: str, driver: ScalarDriver) -> None:
        self._scalar_set(layerId, "setSecondaryFieldLikeTorqueDriver", driver)

    def setLayerSecondaryDampingLikeTorqueDriver(self, layerId: str, driver: ScalarDriver) -> None:
        self._scalar_set(layerId, "setSecondaryDampingLikeTorqueDriver", driver)

    def setLayerPrimaryTorqueDrivers(self, layerId: str, fieldLikeTorque: ScalarDriver,
                                     dampingLikeTorque: ScalarDriver) -> None:
        self._scalar_set(layerId, "setFieldLikeTorqueDriver", fieldLikeTorque)
        self._scalar_set(layerId, "setDampingLikeTorqueDriver", dampingLikeTorque)

    def setLayerSecondaryTorqueDrivers(self, layerId: str, fieldLikeTorque: ScalarDriver,
                                       dampingLikeTorque: ScalarDriver) -> None:
        self._scalar_set(layerId, "setSecondaryFieldLikeTorqueDriver", fieldLikeTorque)
        self._scalar_set(layerId, "setSecondaryDampingLikeTorqueDriver", dampingLikeTorque)

    def setLayerReservedInteractionField(self, layerId: str, driver: AxialDriver) -> None:
        self._scalar_set(layerId, "setReservedInteractionField", driver)

    def setLayerAlternativeSTT(self, layerId: str, alternative: bool) -> None:
        if layerId == "all":
            for l in self.layers:
                l.setAlternativeSTT(alternative)
        else:
            self.getLayer(layerId).setAlternativeSTT(alternative)

    def setLayerOneFNoise(self, layerId: str, sources: int, bias: float, scale: float) -> None:
        targets = self.layers if layerId == "all" else [self.getLayer(layerId)]
        for l in targets:
            l.setOneFNoise(sources, bias, scale)

    def _coupling(self, bottomLayer: str, topLayer: str, driver, top_m: str, bottom_m: str) -> None:
        # setCouplingDriver, core/junction.hpp:1308-1335: adjacent layers only
        for i in range(self.layerNo - 1):
            a, b = self.layers[i].id, self.layers[i + 1].id
            if (a == bottomLayer and b == topLayer) or (a == topLayer and b == bottomLayer):
                getattr(self.layers[i], top_m)(driver)
                getattr(self.layers[i + 1], bottom_m)(driver)
                return
        raise RuntimeError(
            "Failed to match the layer order or find layer ids: " + bottomLayer + " and " + topLayer + "!"
        )

    def setIECDriver(self, bottomLayer: str, topLayer: str, driver: ScalarDriver) -> None:
        self._coupling(bottomLayer, topLayer, driver, "setIECDriverTop", "setIECDriverBottom")

    def setQuadIECDriver(self, bottomLayer: str, topLayer: str, driver: ScalarDriver) -> None:
        self._coupling(bottomLayer, topLayer, driver, "setQuadIECDriverTop", "setQuadIECDriverBottom")

    def setIDMIDriver(self, bottomLayer: str, topLayer: str, driver: AxialDriver) -> None:
        self._coupling(bottomLayer, topLayer, driver, "setIDMIDriverTop", "setIDMIDriverBottom")

    def setLayerMagnetisation(self, layerId: str, mag: CVector) -> None:
        for l in self._each(layerId):
            l.setMagnetisation(mag)

    def getLayerMagnetisation(self, layerId: str) -> CVector:
        return copy.copy(self.getLayer(layerId).mag)

    def getLayerReferenceType(self, layerId: str) -> Reference:
        return self.getLayer(layerId).referenceType

    def setLayerReferenceLayer(self, layerId: str, referenceLayer: CVector) -> None:
        targets = self.layers if layerId == "all" else [self.getLayer(layerId)]
        for l in targets:
            l.setReferenceLayer(_vec(referenceLayer))

    def setLayerReferenceType(self, layerId: str, referenceType: Reference) -> None:
        targets = self.layers if layerId == "all" else [self.getLayer(layerId)]
        for l in targets:
            l.setReferenceLayer(Reference(referenceType))

    def setNoiseSeed(self, seed: int) -> None:
        """Seeds the thermal-noise stream (no reference counterpart: the
        reference seeds from std::random_device and cannot be reproduced)."""
        self._seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self._runs = 0

    # --- magnetoresistance (core/junction.hpp:1835-1889) ---------------------
    def _mr(self, mags) -> list:
        if self.MR_mode == _lib.MR_CLASSIC and self.layerNo == 2:
            c = c_dot(mags[0], mags[1])
            return [self.Rp + (((self.Rap - self.Rp) / 2.0) * (1.0 - c))]
        if self.MR_mode == _lib.MR_CLASSIC and self.layerNo == 1:
            c = c_dot(mags[0], self.layers[0].referenceLayer)
            return [self.Rp + (((self.Rap - self.Rp) / 2.0) * (1.0 - c))]
        if self.MR_mode == _lib.MR_STRIP:
            rx = ry = 0.0
            with np.errstate(divide="ignore"):
                for i, m in enumerate(mags):
                    Rx = self.Rx0[i] + self.AMR_X[i] * (m.x * m.x) + self.SMR_X[i] * (m.y * m.y)
                    Ry = self.Ry0[i] + 0.5 * self.AHE[i] * m.z + (self.AMR_Y[i] + self.SMR_Y[i]) * m.x * m.y
                    rx += float(np.float64(1.0) / np.float64(Rx))
                    ry += float(np.float64(1.0) / np.float64(Ry))
                return [float(np.float64(1) / np.float64(rx)), float(np.float64(1) / np.float64(ry)), 0.0]
        raise RuntimeError("Magnetisation calculation is not supported for this structure!")

    def getMagnetoresistance(self) -> list:
        return self._mr([l.mag for l in self.layers])

    # --- log -------------------------------------------------------------
    def clearLog(self) -> None:
        self._log = {}
        self.logLength = 0

    def getLog(self) -> dict:
        out = {}
        for k, chunks in self._log.items():
            if len(chunks) == 1:
                out[k] = chunks[0].tolist()
            else:
                out[k] = np.concatenate(chunks).tolist()
        return out

    def _append(self, key: str, arr: np.ndarray) -> None:
        # the reference creates a key on its first logged entry: a run that
        # logs nothing leaves the log as it was (logLayerParams, junction.hpp:1580)
        arr = np.asarray(arr, dtype=np.float64)
        if arr.size:
            self._log.setdefault(key, []).append(arr)

    def saveLog(self, filename: str) -> None:
        # saveLogs, core/junction.hpp:1638-1656
        if filename == "":
            raise RuntimeError("The filename may not be empty!")
        log = {k: np.concatenate(v) if v else np.zeros(0) for k, v in self._log.items()}
        with open(filename, "w") as f:
            f.write("".join(f"{k};" for k in log) + "\n")
            for i in range(self.logLength):
                f.write("".join(f"{_fmt(v[i])};" for v in log.values()) + "\n")

    # --- run ---------------------------------------------------------------
    def _solver(self, mode: SolverMode) -> int:
        # getSolver, core/junction.hpp:1891-1931
        local = SolverMode(mode)
        for l in self.layers:
            if l.temperatureSet or l.scaleNoise != 0:
                if local not in (SolverMode.Heun, SolverMode.EulerHeun):
                    print("[WARNING] Solver automatically changed to Euler Heun for stochastic calculation.")
                    local = SolverMode.EulerHeun
            if l.scaleNoise != 0 or l.oneFNoise is not None:
                raise NotImplementedError("coloured (1/f, alpha) noise is not provided by the GPU engine")
        if local == SolverMode.DormandPrince:
            if self.layerNo > 4:
                raise NotImplementedError("DormandPrince runs junctions of 1 to 4 layers on the GPU engine")
        return int(local)

    def _descriptor(self, totalTime: float, timeStep: float, writeFrequency: float, calculateEnergies: bool,
                    solverMode: SolverMode, log_mask: int | None = None):
        """Packs the junction for one run: (junction rec, layer recs, tables, n_steps, W)."""
        if timeStep > writeFrequency:
            raise RuntimeError("The time step cannot be larger than write frequency!")
        n_steps = max(0, int(totalTime / timeStep))
        W = max(1, int(writeFrequency / timeStep))
        solver = self._solver(solverMode)
        for l in self.layers:
            if l.Ms * l.thickness == 0:
                # the reference's IEC term divides a vector by Ms * thickness
                raise RuntimeError("Failed to divide vector by zero!")
            if l.temperatureSet and l.cellVolume == 0.0:
                raise RuntimeError("Cell surface cannot be 0 during temp. calculations!")
        dp = solver == int(SolverMode.DormandPrince)
        if log_mask is None:
            log_mask = _lib.LOG_FULL | (_lib.LOG_ENERGY if calculateEnergies else 0)
        if dp:
            if calculateEnergies:
                raise NotImplementedError("energy channels with DormandPrince are not provided")
            if log_mask:
                log_mask |= _lib.LOG_TIME
        tables = _Tables()
        L = np.zeros(self.layerNo, dtype=_lib.LAYER_DTYPE)
        for i, l in enumerate(self.layers):
            L[i] = l._fill(tables, (n_steps, float(timeStep), solver))
        j = np.zeros((), dtype=_lib.JUNCTION_DTYPE)
        j["n_layers"] = self.layerNo
        j["mr_mode"] = self.MR_mode
        j["solver"] = solver
        j["log_mask"] = log_mask
        j["n_steps"] = n_steps
        j["write_every"] = W
        j["n_entries"] = (n_steps + W - 1) // W
        if dp:
            if tables.array() is not None and len(tables.array()):
                raise NotImplementedError("callback drivers need fixed stage times (not DormandPrince)")
            # the adaptive loop of runSimulation (:1995-2044): at most one entry
            # per write interval before totalTime and one after; attempts capped
            # (at an exact equilibrium the error ratio is 0, the controller's
            # factor saturates at min_factor and the step shrinks to zero
            # before resetting: about 320 attempts per initial step of time)
            j["n_steps"] = min(2**32 - 1, max(1_000_000, 400 * max(1, int(totalTime / timeStep))))
            j["n_entries"] = int(totalTime / writeFrequency) + 4
            j["total_time"] = totalTime
            j["write_freq"] = writeFrequency
        j["seed"] = self._seed
        j["stream"] = self._runs
        j["dt"] = timeStep
        j["Rp"] = self.Rp
        j["Rap"] = self.Rap
        for name in ("Rx0", "Ry0", "AMR_X", "AMR_Y", "SMR_X", "SMR_Y", "AHE"):
            vals = getattr(self, name)
            if vals:
                j[name][: len(vals)] = vals
        return j, L, tables.array(), n_steps, W

    def _apply(self, j, log: np.ndarray, mag: np.ndarray, st, timeStep: float, W: int) -> None:
        """Appends a run's log and takes over its final magnetisation."""
        ne = int(st["entries_done"])
        nl = self.layerNo
        mask = int(j["log_mask"])
        ch = 0
        for l in self.layers:
            for c in ("x", "y", "z"):
                self._append(f"{l.id}_m{c}", log[ch, :ne])
                ch += 1
        if mask & _lib.LOG_MR:
            if self.MR_mode == _lib.MR_CLASSIC:
                self._append(self.Rtag if nl > 1 else "R", log[ch, :ne])
                ch += 1
            elif self.MR_mode == _lib.MR_STRIP:
                self._append("Rx", log[ch, :ne])
                self._append("Ry", log[ch + 1, :ne])
                self._append("Rz", np.zeros(ne))
                ch += 2
        if mask & _lib.LOG_ENERGY:
            for l in self.layers:
                block = log[ch:ch + _lib.ENERGY_CHANNELS, :ne]
                ch += _lib.ENERGY_CHANNELS
                vals = dict(zip(_ENERGY_KEYS, block))
                for k in ("K", "K2", "Jbottom", "Jtop"):
                    self._append(f"{l.id}_{k}", vals[k])
                if l.includeSTT or l.includeSOT:
                    # the reference pushes _I twice per entry (:1593, :1611-1612)
                    self._append(f"{l.id}_I", np.repeat(vals["I"], 2))
                else:
                    self._append(f"{l.id}_I", vals["I"])
                for base in ("Hext", "Hiec", "Hanis", "Hdemag"):
                    for c in ("x", "y", "z"):
                        self._append(f"{l.id}_{base}{c}", vals[base + c])
                for c in ("x", "y", "z"):
                    self._append(f"{l.id}_Hth{c}", np.zeros(ne))
                if l.includeSOT:
                    for base in ("Hfl", "Hdl"):
                        for c in ("x", "y", "z"):
                            self._append(f"{l.id}_{base}{c}", vals[base + c])
        if mask & _lib.LOG_TIME:  # Dormand-Prince: adaptive entry times
            self._append("time", log[ch, :ne])
            ch += 1
        else:
            # entry k is labelled with the pre-step time (k W) dt (core/junction.hpp:1974)
            self._append("time", (np.arange(ne, dtype=np.float64) * W) * timeStep)
        self.logLength += ne
        for i, l in enumerate(self.layers):
            l.mag = CVector(*mag[i])

    def _execute(self, j, L, consts, tables, pid_out=None):
        """One packed run on the GPU through cmtj_run."""
        return _lib.run_junction(j, L, consts, tables, pid_out=pid_out)

    def runSimulation(
        self,
        totalTime: float,
        timeStep: float = 1e-13,
        writeFrequency: float = 1e-11,
        verbose: bool = False,
        calculateEnergies: bool = False,
        solverMode: SolverMode = SolverMode.RK4,
    ) -> None:
        """Integrates the junction on the GPU (reference runSimulation,
        core/junction.hpp:1948-2056, fixed-step solvers)."""
        t0 = time.monotonic()
        j, L, tables, n_steps, W = self._descriptor(totalTime, timeStep, writeFrequency, calculateEnergies,
                                                    solverMode)
        consts = _lib.constants_get()
        pid = None
        if int(j["solver"]) == _lib.DORMAND_PRINCE and np.any(L["dp_use_pid"]):
            # the PID controller's integral and last error stay in the layer
            # from run to run (AdaptiveIntegrationParams, core/junction.hpp:953-961)
            pid = np.stack([L["dp_integral_error"], L["dp_prev_error_ratio"]], axis=1)
            log, mag, st = self._execute(j, L, consts, tables, pid_out=pid)
        else:
            log, mag, st = self._execute(j, L, consts, tables)
        self._runs += 1
        self._apply(j, log, mag, st, timeStep, W)
        if pid is not None:
            for l, (integral, prev) in zip(self.layers, pid):
                l.adaptiveParams.integral_error = float(integral)
                l.adaptiveParams.prev_error_ratio = float(prev)
        if int(st["code"]) == _lib.ERR_NAN:
            raise RuntimeError("NAN magnetisation")
        if int(st["code"]) == _lib.ERR_LIMIT:
            raise RuntimeError("DormandPrince: the step-attempt or log-entry limit of the run was reached")
        if int(st["code"]) == _lib.ERR_DP_REJECTED:
            # runMultiLayerSolver (core/junction.hpp:1699-1704) takes the
            # smallest Layer::hopt after a rejection; hopt stays -1, so the
            # reference's clock runs backwards from here and never finishes
            raise RuntimeError("DormandPrince: a layer rejected its step; the reference then sets the step to "
                               "Layer::hopt = -1 and its clock runs backwards (the run cannot finish). "
                               "Start from a smaller timeStep or loosen the tolerances")
        if int(st["code"]) != _lib.OK:
            raise RuntimeError(f"engine status {int(st['code'])}")
        if verbose:
            print(f"Steps in simulation: {n_steps}")
            print(f"Write every: {W}")
            print(f"Simulation time = {int(time.monotonic() - t0)}[s]")


__all__ = [
    "Axis", "Reference", "SolverMode", "UpdateType", "CVector", "c_dot", "ScalarDriver", "NullDriver",
    "AdaptiveIntegrationParams",
    "AxialDriver", "Layer", "Junction", "constantDriver", "pulseDriver", "sineDriver", "posSineDriver",
    "stepDriver", "trapezoidDriver", "gaussianImpulseDriver", "gaussianStepDriver",
]
