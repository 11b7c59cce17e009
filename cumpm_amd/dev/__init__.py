from .electric_field import MonodisperseEwaldElectricField, PolydisperseEwaldElectricField, FieldCalcMode

__all__ = ["MonodisperseEwaldElectricField", "PolydisperseEwaldElectricField", "FieldCalcMode"]
