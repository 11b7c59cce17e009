from .electric_field import MonodisperseEwaldElectricField, FieldCalcMode

__all__ = ["MonodisperseEwaldElectricField", "FieldCalcMode"]
