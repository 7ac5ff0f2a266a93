"""algotom_amd: algotom's filtered back-projection path, MI355X-native.

    from algotom_amd.rec import reconstruction as rec
    rec.fbp_reconstruction(sinogram, center, ...)
"""

__version__ = "0.1.0"
