import os
import sys

from .emit import generate

TARGET = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "csrc", "generated",
                      "elasticity_iso.inc")

if __name__ == "__main__":
    txt = generate()
    if "--check" in sys.argv:
        with open(TARGET) as f:
            if f.read() != txt:
                print("stale: %s differs from the generator output" % TARGET)
                sys.exit(1)
        print("up to date")
        sys.exit(0)
    os.makedirs(os.path.dirname(TARGET), exist_ok=True)
    with open(TARGET, "w") as f:
        f.write(txt)
    print("wrote %s" % TARGET)
