"""XML parameter files: utils/Conf.java's typed folder tree.

A document is a tree of folders (XML elements without a ``type``
attribute) holding typed entries (elements with ``type`` = int, decimal,
string, data or timepoint).  Text forms follow Conf.java:
int " 1 2 ", decimal " %8.5e " per value (or " BASE64:...:END64 " of
big-endian doubles when written exact), data " BASE64:...:END64 " (standard
alphabet, padded), timepoint " <ms> ".  Numbers are read as Java's Scanner
reads them: whitespace-separated tokens up to the first one that does not
parse.  Binary float payloads (OTF bands) are big-endian float32, as
Conf.fromByte / toByte use Java's default ByteBuffer order
(Conf.java:759-779).
"""
import base64
import math
import struct
import xml.etree.ElementTree as ET


class EntryNotFound(KeyError):
    """Conf.EntryNotFoundException."""


def _java_double(v):
    """String.format("%8.5e", v) with Java's spellings of the specials."""
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "Infinity" if v > 0 else "-Infinity"
    return "%8.5e" % v


def _scan(text, conv):
    out = []
    for tok in text.split():
        try:
            out.append(conv(tok))
        except ValueError:
            break
    return out


def _jfloat(tok):
    if tok in ("NaN", "Infinity", "-Infinity", "+Infinity"):
        return float(tok.replace("Infinity", "inf"))
    if tok.lower() in ("nan", "inf", "-inf", "+inf", "infinity", "-infinity"):
        raise ValueError(tok)
    return float(tok)


def _b64_payload(text):
    i, j = text.index("BASE64:") + 7, text.index(":END64")
    return base64.b64decode(text[i:j].strip())


class Entry:
    type = None


class IntEntry(Entry):
    type = "int"

    def __init__(self, vals=(0,)):
        self.vals = [int(v) for v in vals]

    def val(self):
        return self.vals[0]

    def text(self):
        return " " + "".join(f"{v} " for v in self.vals)

    def from_text(self, t):
        self.vals = _scan(t, int)


class DoubleEntry(Entry):
    type = "decimal"

    def __init__(self, vals=(0.0,), exact=False):
        self.vals = [float(v) for v in vals]
        self.exact = exact

    def val(self):
        return self.vals[0]

    def text(self):
        if not self.exact:
            return " " + "".join(_java_double(v) + " " for v in self.vals)
        raw = struct.pack(">%dd" % len(self.vals), *self.vals)
        return " BASE64:" + base64.b64encode(raw).decode() + ":END64 "

    def from_text(self, t):
        if "BASE64:" in t:
            raw = base64.b64decode(t[t.index(":") + 1:t.index(":END64")].strip())
            if len(raw) % 8:
                raise ValueError("Wrong base64 len for double")
            self.vals = list(struct.unpack(">%dd" % (len(raw) // 8), raw))
        else:
            self.vals = _scan(t, _jfloat)


class StringEntry(Entry):
    type = "string"

    def __init__(self, v=""):
        self.v = v

    def val(self):
        return self.v

    def text(self):
        return self.v

    def from_text(self, t):
        self.v = t


class DataEntry(Entry):
    type = "data"

    def __init__(self, data=b""):
        self.data = bytes(data)

    def val(self):
        return self.data

    def text(self):
        return " BASE64:" + base64.b64encode(self.data).decode() + ":END64 "

    def from_text(self, t):
        self.data = _b64_payload(t)


class TimeDateEntry(Entry):
    type = "timepoint"

    def __init__(self, ms=0):
        self.ms = int(ms)

    def val(self):
        return self.ms

    def text(self):
        return f" {self.ms} "

    def from_text(self, t):
        v = _scan(t, int)
        if v:
            self.ms = v[0]


_TYPES = {c.type: c for c in (IntEntry, DoubleEntry, StringEntry, DataEntry, TimeDateEntry)}


class Folder(Entry):
    """Conf.Folder: named entries and sub-folders."""

    def __init__(self, name):
        self.name = name
        self.entries = {}

    def contains(self, name):
        return name in self.entries

    def _get(self, name, cls):
        e = self.entries.get(name)
        if not isinstance(e, cls):
            raise EntryNotFound(name)
        return e

    def cd(self, name):
        return self._get(name, Folder)

    def mk(self, name):
        e = self.entries.get(name)
        if not isinstance(e, Folder):
            e = Folder(name)
            self.entries[name] = e
        return e

    def folders(self):
        return [e for e in self.entries.values() if isinstance(e, Folder)]

    def getInt(self, name):
        return self._get(name, IntEntry)

    def getDbl(self, name):
        return self._get(name, DoubleEntry)

    def getStr(self, name):
        return self._get(name, StringEntry)

    def getData(self, name):
        return self._get(name, DataEntry)

    def getTDE(self, name):
        return self._get(name, TimeDateEntry)

    def newInt(self, name, *vals):
        self.entries[name] = IntEntry(vals or (0,))
        return self.entries[name]

    def newDbl(self, name, *vals, exact=False):
        self.entries[name] = DoubleEntry(vals or (0.0,), exact)
        return self.entries[name]

    def newStr(self, name, v=""):
        self.entries[name] = StringEntry(v)
        return self.entries[name]

    def newData(self, name, data=b""):
        self.entries[name] = DataEntry(data)
        return self.entries[name]

    def newTDE(self, name, ms=0):
        self.entries[name] = TimeDateEntry(ms)
        return self.entries[name]

    def to_element(self):
        el = ET.Element(self.name)
        for n, e in self.entries.items():
            if isinstance(e, Folder):
                el.append(e.to_element())
            else:
                sub = ET.SubElement(el, n, {"type": e.type})
                sub.text = e.text()
        return el


class Conf:
    """Conf.java: a root folder named after the namespace."""

    def __init__(self, namespace):
        self.namespace = namespace
        self.root = Folder(namespace)

    def r(self):
        return self.root

    def cd(self, namespace):
        if namespace != self.namespace:
            raise EntryNotFound(namespace)
        return self.root

    @staticmethod
    def _import(el, fld):
        """Conf.importXmlElement (Conf.java:682-731)."""
        for e in el:
            if not isinstance(e.tag, str):
                continue   # comments, processing instructions
            t = e.get("type", "")
            if t == "":
                Conf._import(e, fld.mk(e.tag))
                continue
            cls = _TYPES.get(t)
            if cls is not None:
                ent = cls()
                ent.from_text("".join(e.itertext()))
                fld.entries[e.tag] = ent

    @staticmethod
    def loadFile(path):
        """Conf.loadFile (Conf.java:638-677)."""
        root = ET.parse(path).getroot()
        cfg = Conf(root.tag)
        Conf._import(root, cfg.root)
        return cfg

    def saveFile(self, path):
        """Conf.saveFile (Conf.java:597-632): indented XML, no declaration."""
        el = self.root.to_element()
        ET.indent(el)
        ET.ElementTree(el).write(path, encoding="utf-8", xml_declaration=False)
        return True


def from_byte(data):
    """Conf.fromByte: big-endian float32."""
    if len(data) % 4:
        raise ValueError("Array size not a multiple of 4")
    return list(struct.unpack(">%df" % (len(data) // 4), data))


def to_byte(vals):
    """Conf.toByte: big-endian float32."""
    return struct.pack(">%df" % len(vals), *vals)
