"""neuralat_amd: MI355X-native BEM Helmholtz path of NeuralAT.

``neuralat_amd.bem`` mirrors the reference's ``src/bem`` API over the HIP
library ``libbem_hip.so`` (C ABI: include/bem_hip.h); ``neuralat_amd.mesh``
builds synthetic meshes and analytic fields.
"""
__version__ = "0.1.0"
