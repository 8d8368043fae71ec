"""Copies the reference's real-data files used by the event-I/O tests into
tests/golden/cms/ (data files, not code):

  plotting/data/physics/hits.ttBar.pb   10 CMS ttbar events (PEventContainer)
  configs/detectorRadius.dat            detector id, radius-dictionary entry, layer
  configs/radiusDictionary.dat          dictionary entry, module radius (cm)

Run from the repository root in a container where the reference is mounted:
    python tests/golden/make_cms_fixtures.py
"""
import os
import shutil

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cms")

FILES = {
    "hits.ttBar.pb": "plotting/data/physics/hits.ttBar.pb",
    "detectorRadius.dat": "configs/detectorRadius.dat",
    "radiusDictionary.dat": "configs/radiusDictionary.dat",
}

if __name__ == "__main__":
    os.makedirs(OUT, exist_ok=True)
    for dst, src in FILES.items():
        shutil.copyfile(os.path.join(REF, src), os.path.join(OUT, dst))
        print("copied", src)
